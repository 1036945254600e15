"""HIP-graph capture of a whole q-point step (zmel, x0gemm, Hilbert, W build,
Sigma_c) against eager execution.  Each replay is checked against the eager
result, with eager work and host synchronisation between the replays; the
graph holds no memset nodes (memory is cleared by fill kernels)."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_qpoint_graph_replays_match_eager(prec):
    from ecalj_amd.qsgw import QPointGPU

    say("graph", prec)
    qp = synth.make_qpoint(synth.CONFIGS["T3"])
    g = QPointGPU(qp, prec)
    g.step()
    g.make_wk()
    g.step()                      # warm-up at the final shapes (scratch sized, tuner settled)
    chi0, wc0, zsec0, eqp0 = g.results()
    graph = g.capture_step()
    try:
        for rep in range(3):
            g.chi.zero()          # eager work between replays; the replay must recompute everything
            g.zsec.zero()
            g.h.sync()
            graph.launch()
            chi, wc, zsec, eqp = g.results()
            say("replay", rep, np.abs(zsec - zsec0).max())
            assert np.array_equal(chi, chi0)
            assert np.array_equal(wc, wc0)
            assert np.array_equal(zsec, zsec0)
            g.step()              # an eager step in between gives the same numbers too
            assert np.array_equal(g.results()[2], zsec0)
    finally:
        graph.close()
        g.h.close()


def test_capture_without_warmup_is_refused():
    """Scratch that would have to grow under capture is refused with a state
    error (nothing is launched); the handle then works eagerly as before."""
    import ecalj_amd
    from ecalj_amd.qsgw import QPointGPU

    qp = synth.make_qpoint(synth.CONFIGS["T1"])
    g = QPointGPU(qp, "f64")
    try:
        g.h.graph_begin()
        with pytest.raises(ecalj_amd.EcaljError) as e:
            g.chi0()
        assert e.value.status == 6  # ECALJ_ERR_STATE
        g.h.graph_end().close()
        g.step()
        assert np.isfinite(g.results()[3]).all()
    finally:
        g.h.close()
