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


def test_graph_refused_after_scratch_moves():
    """A graph keeps the scratch addresses of its capture.  An eager call that
    makes the same handle's scratch grow (here a larger zmel) frees the old
    scratch, so a replay of the older graph is refused with a state error on
    the host instead of running on freed memory; so is a replay after
    set_coulomb, and a launch on a handle other than the capturing one.  A
    graph recorded after the growth replays normally."""
    import ctypes as C

    import ecalj_amd
    from ecalj_amd.qsgw import QPointGPU

    qp = synth.make_qpoint(synth.CONFIGS["T3"])
    g = QPointGPU(qp, "f64")
    other = ecalj_amd.Handle(0, "f64")
    try:
        g.step()
        g.make_wk()
        g.step()
        zsec0 = g.results()[2]
        g1 = g.capture_step()
        g1.launch()
        assert np.array_equal(g.results()[2], zsec0)
        st = ecalj_amd.lib().ecalj_graph_launch(other.ptr, C.c_void_p(g1.ptr))
        assert st == 4  # ECALJ_ERR_ARG: the graph belongs to another handle
        say("graph: growing the handle's scratch with a larger eager zmel")
        rng = np.random.default_rng(1)
        nb = qp.cfg.nb
        # both sides wide, so the contraction's intermediate (ndima x nI x 300) outgrows the step's scratch
        big = rng.standard_normal((nb, 300)) + 1j * rng.standard_normal((nb, 300))
        g.h.zmel(big, big, zmelconjg=True)
        with pytest.raises(ecalj_amd.EcaljError) as e:
            g1.launch()
        assert e.value.status == 6  # ECALJ_ERR_STATE
        g.step()                    # eager work is unaffected
        assert np.array_equal(g.results()[2], zsec0)
        g2 = g.capture_step()       # recorded at the new generation: replays
        g.zsec.zero()
        g2.launch()
        assert np.array_equal(g.results()[2], zsec0)
        g.h.set_coulomb(qp.vcoud, qp.zcousq)
        with pytest.raises(ecalj_amd.EcaljError) as e:
            g2.launch()
        assert e.value.status == 6
        g1.close()
        g2.close()
    finally:
        other.close()
        g.h.close()


def test_capture_refused_under_the_profiler(monkeypatch):
    """Under rocprofv3 (its ROCPROF_* variables set) graph capture is refused
    with a state error before anything is recorded, and eager calls run."""
    import ecalj_amd
    from ecalj_amd.qsgw import QPointGPU

    qp = synth.make_qpoint(synth.CONFIGS["T1"])
    g = QPointGPU(qp, "f64")
    try:
        monkeypatch.setenv("ROCPROF_KERNEL_TRACE", "1")
        with pytest.raises(ecalj_amd.EcaljError) as e:
            g.h.graph_begin()
        assert e.value.status == 6  # ECALJ_ERR_STATE
        g.step()
        assert np.isfinite(g.results()[3]).all()
        monkeypatch.delenv("ROCPROF_KERNEL_TRACE")
        g.h.graph_begin()
        g.step()
        g.h.graph_end().close()
    finally:
        g.h.close()
