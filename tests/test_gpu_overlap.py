"""Two streams within one q-point: Sigma_c's zmel and weights on a second
handle beside the W build (QPointGPU(overlap=True)), joined by
ecalj_stream_wait before the contraction.  The same kernels run on the same
inputs (the Sigma_c kernel variant fixed), so every result equals the
single-stream one exactly, also over repeated steps (the second handle waits for the previous contraction before
it overwrites zmel)."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_overlap_equals_single_stream(prec):
    from ecalj_amd.qsgw import QPointGPU

    print("overlap", prec, flush=True)
    sys.stdout.flush()
    qp = synth.make_qpoint(synth.CONFIGS["T3"])
    out = []
    for ov in (False, True):
        g = QPointGPU(qp, prec, overlap=ov)
        # one kernel variant for both runs: the tuner times the variants anew
        # on every handle, and variants of different forms round differently
        for hd in (g.h, getattr(g, "h2", None)):
            if hd is not None:
                hd.set_tuning(10)
        try:
            g.step()
            g.make_wk()
            for _ in range(3):
                g.step()
            chi, wc, zsec, eqp = g.results()
        finally:
            g.close()
        out.append((chi, wc, zsec, eqp))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, b)


def test_stream_wait_rejects_capture():
    import ecalj_amd

    print("stream_wait under capture", flush=True)
    h1, h2 = ecalj_amd.Handle(0, "f64"), ecalj_amd.Handle(0, "f64")
    try:
        x = h1.array((64,))
        h1.stream_wait(h2)
        h1.graph_begin()
        try:
            x.zero()  # one kernel node
            with pytest.raises(ecalj_amd.EcaljError):
                h1.stream_wait(h2)
        finally:
            h1.graph_end().close()
        h1.sync()
    finally:
        h1.close()
        h2.close()
