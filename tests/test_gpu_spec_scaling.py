"""The workloads under their literal ppb statement, N(0, 1/nb), instead of the
cell-normalised variance the other tests use (natom^2/(nI*Omega)).

With 1/nb, sum_I |zmel|^2 = nI/nb: chi0 grows ~1000x, epsilon reaches ~10^4 and
Sigma_c ~10^3 Ry, so the QP shifts are thousands of eV.  The FP64 form keeps
its 1e-10 bar; the mixed forms keep their per-result relative bars (chi0,
W_c, Sigma_c), and their QP-energy differences are printed beside the largest
|Sigma_c| they sit on.  A 1 meV bar on a 10^4 eV self-energy is a relative
bar of 1e-7, below what FP32 accumulation holds, so it is reported, not
asserted, in this scaling.  C2 shape (nI = 600, nb = 180, 9 + 150 bands,
nf = 16) on 4 of its 64 k-points."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-5, "bf16x3": 1e-4}
SIG_TOL = {"f32": 2e-4, "bf16x3": 1e-3}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def c2s():
    from oracle import pipeline

    qp = synth.make_qpoint(synth.CONFIGS["C2"].scaled(nk=4, ppb_var="spec"))
    say("oracle: C2 nk=4, ppb ~ N(0, 1/nb) ...")
    ref = pipeline.run(qp)
    say("oracle: done")
    return qp, ref


def _run(qp, prec):
    from ecalj_amd.qsgw import QPointGPU

    say(f"gpu: C2 spec scaling {prec}")
    g = QPointGPU(qp, prec)
    try:
        g.chi0()
        g.wbuild()
        g.make_wk()
        g.sigc()
        return g.results()
    finally:
        g.h.close()


def test_spec_scaling_f64(c2s):
    qp, ref = c2s
    chi, wc, zsec, eqp = _run(qp, "f64")
    errs = rel(chi, ref.chi), rel(wc, ref.wc), rel(zsec, ref.zsec)
    say(f"  f64: chi {errs[0]:.2e} W {errs[1]:.2e} Sigma {errs[2]:.2e} "
        f"|dEqp| {np.abs(eqp - ref.eqp).max() * 1e3:.3e} meV")
    assert max(errs) < 1e-10
    assert np.abs(eqp - ref.eqp).max() < 1e-7 * np.abs(ref.eqp).max()


@pytest.mark.parametrize("prec", ["bf16x3", "f32"])
def test_spec_scaling_mixed(c2s, prec):
    from ecalj_amd.synth import RYDBERG_EV

    qp, ref = c2s
    chi, wc, zsec, eqp = _run(qp, prec)
    errs = rel(chi, ref.chi), rel(wc, ref.wc), rel(zsec, ref.zsec)
    mev = float(np.abs(eqp - ref.eqp).max() * 1e3)
    sig_ev = float(np.abs(ref.zsec).max() * RYDBERG_EV)
    say(f"  {prec}: chi {errs[0]:.2e} W {errs[1]:.2e} Sigma {errs[2]:.2e} |dEqp| {mev:.3f} meV "
        f"on max|Sigma_c| {sig_ev:.1f} eV (relative {mev * 1e-3 / sig_ev:.2e})")
    assert errs[0] < TOL[prec]
    assert errs[1] < 10 * TOL[prec]
    assert errs[2] < SIG_TOL[prec]
