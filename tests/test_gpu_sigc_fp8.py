"""Sigma_c frequency contraction with fp8 cross terms (precision "bf16f8":
everything as bf16x3 except that in the Sigma_c contraction each product
a*b = (ah + al)(bh + bl) is ah*bh on the bf16 MFMA plus ah*bl + al*bh on the
scaled fp8 MFMA, E4M3 operands with one E8M0 scale per 32 values).

Bars, stated per result as max|gpu - oracle| / max|oracle|:
  Sigma_c : 1e-3.  E4M3 keeps 3 mantissa bits, so each rounded cross-term
            operand is off by at most 2^-4 of itself; |ah*bl|, |al*bh| <=
            2^-9 |ab|, so one product is off by at most ~2 * 2 * 2^-4 * 2^-9
            = 2^-11 ~ 4.9e-4 of |ab| (the dropped al*bl adds 2^-18); the bar
            is twice that worst case.
  QP energies of whole q-points: 1 meV, the accuracy bar of the mixed forms.
The other contractions of this precision are the bf16x3 ones (1e-4)."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu
TOL_SIGC = 1e-3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def crand(rng, *shape):
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def relerr(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def h8():
    import ecalj_amd

    hd = ecalj_amd.Handle(0, "bf16f8")
    yield hd
    hd.close()


# ragged, tile-aligned, more than 32 frequencies (two launches), a single row
@pytest.mark.parametrize("ngb,nm,ntq,niw", [(130, 21, 8, 12), (256, 64, 2, 10), (161, 13, 5, 40), (1, 3, 1, 2)])
def test_sigc_fp8(h8, orc, ngb, nm, ntq, niw):
    say("sigc bf16f8", ngb, nm, ntq, niw)
    rng = np.random.default_rng(ngb + 1)
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    ref = orc.sigc_imag(z, wc, g)
    err = relerr(h8.sigc(z, wc, g), ref)
    say(f"  relerr {err:.2e}")
    assert err < TOL_SIGC


def test_sigc_fp8_dynamic_range(h8, orc):
    """Rows, columns and frequencies of W over 10^-6..10^2 and zmel columns
    over 10^-3..10^3: each 32-value block gets its own scale, so no block
    under- or overflows E4M3; a single global scale would flush the small
    blocks (E4M3 spans 2^-9..448)."""
    rng = np.random.default_rng(77)
    ngb, nm, ntq, niw = 150, 17, 6, 9
    d = 10.0 ** rng.uniform(-3, 1, ngb)
    f = 10.0 ** np.linspace(0, -4, niw + 1)
    wc = np.asfortranarray(crand(rng, ngb, ngb, niw + 1) * d[:, None, None] * d[None, :, None] * f[None, None, :])
    z = np.asfortranarray(crand(rng, ngb, nm, ntq) * (10.0 ** rng.uniform(-3, 3, (1, nm, ntq))))
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    ref = orc.sigc_imag(z, wc, g)
    err = relerr(h8.sigc(z, wc, g), ref)
    say(f"sigc bf16f8 dynamic range relerr {err:.2e}")
    assert err < TOL_SIGC


def test_sigc_fp8_tracks_bf16x3(orc):
    """Against the oracle the fp8 cross terms sit between bf16x3 and plain
    bf16 (one product per term, ~2^-9): better than bf16 rounding of W, worse
    than bf16x3."""
    import ecalj_amd

    rng = np.random.default_rng(5)
    ngb, nm, ntq, niw = 200, 19, 7, 11
    z = crand(rng, ngb, nm, ntq)
    wc = crand(rng, ngb, ngb, niw + 1)
    g = np.asfortranarray(rng.standard_normal((niw + 1, nm, ntq)))
    ref = orc.sigc_imag(z, wc, g)
    errs = {}
    for prec in ("bf16x3", "bf16f8"):
        hd = ecalj_amd.Handle(0, prec)
        try:
            errs[prec] = relerr(hd.sigc(z, wc, g), ref)
        finally:
            hd.close()
    # plain bf16 rounding of W alone (the oracle on the rounded W)
    def bf16(x):
        u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        return u.astype(np.uint32).view(np.float32).astype(np.float64)
    wb = np.asfortranarray(bf16(wc.real) + 1j * bf16(wc.imag))
    errs["bf16 W"] = relerr(orc.sigc_imag(z, wb, g), ref)
    say("relerr", {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["bf16x3"] < errs["bf16f8"] < errs["bf16 W"]


def test_sigc_fp8_grouped_equals_uniform():
    """Same windows at every k: the grouped persistent kernel and the uniform
    batch run the same tile code in the same order."""
    import ecalj_amd

    rng = np.random.default_rng(9)
    ngb, nm, ntq, niw, nk = 72, 11, 4, 6, 3
    h = ecalj_amd.Handle(0, "bf16f8")
    try:
        z = (rng.standard_normal((ngb, nm, ntq, nk)) + 1j * rng.standard_normal((ngb, nm, ntq, nk))).astype(
            np.complex64)
        a = rng.standard_normal((ngb, ngb, niw + 1, nk)) + 1j * rng.standard_normal((ngb, ngb, niw + 1, nk))
        w = a + a.conj().transpose(1, 0, 2, 3)
        g = rng.standard_normal((niw + 1, nm, ntq, nk))
        dz, dw, dg = h.to_dev(z), h.to_dev(w), h.to_dev(g)
        s1, s2 = h.array((ntq, ntq)), h.array((ntq, ntq))
        s1.zero()
        s2.zero()
        h.sigc_dev(nk, dz, ngb, nm, ntq, niw, dw, dg, s1)
        h.sigc_grouped_dev([nm] * nk, dz, ngb, ntq, niw, dw, dg, s2)
        a, b = s1.download(), s2.download()
    finally:
        h.close()
    say("grouped vs uniform", relerr(b, a))
    assert relerr(b, a) < 1e-6


@pytest.mark.parametrize("name", ["T3", "C1"])
def test_qpoint_fp8_qp_energy_within_1meV(orc, name):
    from ecalj_amd.qsgw import QPointGPU
    from oracle import pipeline

    say("qpoint bf16f8", name)
    qp = synth.make_qpoint(synth.CONFIGS[name])
    ref = pipeline.run(qp)
    g = QPointGPU(qp, "bf16f8")
    g.chi0()
    g.wbuild()
    g.make_wk()
    g.sigc()
    chi, wc, zsec, eqp = g.results()
    g.close()
    err_mev = np.abs(eqp - ref.eqp).max() * 1e3
    say(f"  max |dEqp| = {err_mev:.4f} meV, Sigma relerr {relerr(zsec, ref.zsec):.2e}")
    assert err_mev < 1.0
    assert relerr(chi, ref.chi) < 1e-4


def test_uneven_fp8_grouped():
    from ecalj_amd.batched import UnevenQPointGPU
    from oracle import pipeline

    cfg = synth.Config("U2", nk=5, nb=96, no=9, nu=70, ni=136, nf=10, seed=32, natom=4)
    qp = synth.make_uneven_qpoint(cfg, no_range=(3, 9), nu_range=(20, 70))
    ref = pipeline.run(qp)
    say("uneven bf16f8")
    g = UnevenQPointGPU(qp, "bf16f8")
    g.chi0()
    g.wbuild()
    g.make_wk()
    g.sigc()
    _, _, zsec, eqp = g.results()
    g.h.close()
    say(f"  Sigma relerr {relerr(zsec, ref.zsec):.2e}, max |dEqp| {np.abs(eqp - ref.eqp).max() * 1e3:.4f} meV")
    assert relerr(zsec, ref.zsec) < TOL_SIGC
    assert np.abs(eqp - ref.eqp).max() < 1.0e-3  # eV
