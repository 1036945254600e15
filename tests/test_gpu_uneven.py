"""q-points with uneven per-k band windows (the batched workload's shape):
per-k zmel, x0gemm over per-k occupied ranges, and Sigma_c as one grouped
persistent kernel (ecalj_sigc_grouped_dev), against the FP64 oracle
pipeline with the same windows.  Bars as in the uniform tests: FP64 1e-10
relative; bf16x3 1e-4 relative on Sigma_c and 1 meV on QP energies."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu

U1 = synth.Config("U1", nk=3, nb=48, no=6, nu=30, ni=64, nf=6, seed=31, natom=2)
U2 = synth.Config("U2", nk=5, nb=96, no=9, nu=70, ni=136, nf=10, seed=32, natom=4)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module", params=[("U1", (2, 6), (10, 30)), ("U2", (3, 9), (20, 70))])
def case(request):
    from oracle import pipeline

    name, nr, ur = request.param
    cfg = {"U1": U1, "U2": U2}[name]
    qp = synth.make_uneven_qpoint(cfg, no_range=nr, nu_range=ur)
    say("uneven", name, "no_k", list(qp.extra["no_k"]), "nu_k", list(qp.extra["nu_k"]))
    return qp, pipeline.run(qp)


def test_uneven_f64_parity(case):
    from ecalj_amd.batched import UnevenQPointGPU

    qp, ref = case
    g = UnevenQPointGPU(qp, "f64")
    g.chi0()
    g.wbuild()
    g.make_wk()     # W_c at the internal k from this W_c(q), as the oracle pipeline does
    g.sigc()
    chi, wc, zsec, eqp = g.results()
    g.h.close()
    assert rel(chi, ref.chi) < 1e-10
    assert rel(wc, ref.wc) < 1e-10
    assert rel(zsec, ref.zsec) < 1e-10
    assert np.abs(eqp - ref.eqp).max() < 1e-8


@pytest.mark.parametrize("mode", [1, 2, 4, 5])
def test_uneven_bf16x3_grouped(case, mode):
    """Each tile variant of the grouped persistent kernel, forced by the tuner
    (4, 5: the three-product form in its 2 x 4 and 1 x 8 wave layouts, whose
    tiles are also the uniform frequency-inner kernels')."""
    from ecalj_amd.batched import UnevenQPointGPU

    qp, ref = case
    g = UnevenQPointGPU(qp, "bf16x3")
    g.h.set_tuning(mode)
    g.chi0()
    g.wbuild()
    g.make_wk()
    g.sigc()
    _, _, zsec, eqp = g.results()
    g.h.close()
    assert rel(zsec, ref.zsec) < 1e-4
    assert np.abs(eqp - ref.eqp).max() < 1.0e-3  # eV


def test_grouped_equals_uniform_when_windows_agree():
    """With every k at the same window, the grouped entry point gives the batched one's Sigma_c."""
    import ecalj_amd

    rng = np.random.default_rng(5)
    ngb, nm, ntq, niw, nk = 72, 11, 4, 6, 3
    # (precision, tuner mode of the uniform kernel, of the grouped one, bar).  The
    # forced pairs run the same arithmetic (four products, 128x64 tiles; three
    # products); the tuner's own choices may pair a three-product kernel with a
    # four-product one, whose roundings differ by a few bf16x3 ulps of these
    # random, cancelling sums.
    for prec, mu, mg, tol in (("f64", 0, 0, 1e-13), ("bf16x3", 1, 1, 1e-6), ("bf16x3", 5, 4, 1e-6),
                              ("bf16x3", 0, 0, 2e-5)):
        h = ecalj_amd.Handle(0, prec)
        cst = np.complex128 if prec == "f64" else np.complex64
        z = (rng.standard_normal((ngb, nm, ntq, nk)) + 1j * rng.standard_normal((ngb, nm, ntq, nk))).astype(cst)
        a = rng.standard_normal((ngb, ngb, niw + 1, nk)) + 1j * rng.standard_normal((ngb, ngb, niw + 1, nk))
        w = a + a.conj().transpose(1, 0, 2, 3)  # Hermitian, as W_c(i nu) is: clipping per k or once then agree
        g = rng.standard_normal((niw + 1, nm, ntq, nk))
        dz, dw, dg = h.to_dev(z), h.to_dev(w), h.to_dev(g)
        s1, s2 = h.array((ntq, ntq)), h.array((ntq, ntq))
        s1.zero()
        s2.zero()
        h.set_tuning(mu)
        h.sigc_dev(nk, dz, ngb, nm, ntq, niw, dw, dg, s1)
        h.set_tuning(mg)
        h.sigc_grouped_dev([nm] * nk, dz, ngb, ntq, niw, dw, dg, s2)
        a, b = s1.download(), s2.download()
        h.close()
        say(prec, mu, mg, rel(b, a))
        assert rel(b, a) < tol


def test_uneven_graph_replay(case):
    from ecalj_amd.batched import UnevenQPointGPU

    qp, _ = case
    g = UnevenQPointGPU(qp, "bf16x3")
    g.step()
    g.make_wk()
    g.step()
    _, _, z0, _ = g.results()
    gr = g.capture_step()
    try:
        for _ in range(2):
            g.zsec.zero()
            g.h.sync()
            gr.launch()
            assert np.array_equal(g.results()[2], z0)
    finally:
        gr.close()
        g.h.close()
