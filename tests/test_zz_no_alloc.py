"""ecalj_zminv_noalloc_dev, the entry point that allocates nothing, against
ecalj_zminv_dev (same kernels, so the results must be bit-identical) through
the C ABI: two calls in a row on one handle, the second smaller; a request
beyond the fixed scratch refused before any work; and a first use inside a
HIP graph capture with no warm-up call (any allocation there would be refused
by the library), replayed several times with eager work between replays."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _matrices(rng, n, lda, nb):
    """nb matrices (lda x n, padding rows 7+7i) whose in-order elimination meets a zero pivot."""
    a = np.zeros((lda, n, nb), complex, order="F")
    for b in range(nb):
        x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        x[0, 0] = 0.0
        x[2, 2] = 1e-13
        a[:n, :, b] = x
        a[n:, :, b] = 7.0 + 7.0j
    return a


def _both(h, a, n, lda, nb):
    d1 = h.to_dev(a)
    d2 = h.to_dev(a)
    h.zminv_dev(d1, n, nbatch=nb, lda=lda)
    h.zminv_noalloc_dev(d2, n, nbatch=nb, lda=lda)
    h.sync()
    g1 = d1.download().reshape(a.shape, order="F")
    g2 = d2.download().reshape(a.shape, order="F")
    d1.free()
    d2.free()
    return g1, g2


@pytest.mark.parametrize("first,second", [((600, 605, 3), (100, 100, 2))])
def test_noalloc_matches_zminv_dev(first, second):
    import ecalj_amd

    h = ecalj_amd.Handle(0, "f64")
    try:
        rng = np.random.default_rng(355)
        for n, lda, nb in (first, second):
            say(f"zminv_noalloc_dev vs zminv_dev: n={n} lda={lda} nbatch={nb}")
            a = _matrices(rng, n, lda, nb)
            g1, g2 = _both(h, a, n, lda, nb)
            assert np.array_equal(g1, g2)  # same kernels, same order
            for b in range(nb):
                ref = np.linalg.inv(a[:n, :, b])
                assert np.abs(g2[:n, :, b] - ref).max() / np.abs(ref).max() < 1e-9
            assert np.all(g2[n:] == 7.0 + 7.0j)
    finally:
        h.close()


def test_noalloc_refuses_beyond_scratch():
    import ecalj_amd

    h = ecalj_amd.Handle(0, "f64")
    try:
        n = 1900  # 16 n^2 alone is 57.8 MB; with the panels it exceeds 64 MiB
        say(f"zminv_noalloc_dev refusal: n={n}")
        a = np.asfortranarray(np.eye(n, dtype=complex) * 2.0)
        d = h.to_dev(a)
        with pytest.raises(ecalj_amd.EcaljError) as e:
            h.zminv_noalloc_dev(d, n)
        assert e.value.status == 9
        h.sync()
        assert np.array_equal(d.download().reshape(a.shape, order="F"), a)  # nothing was queued
        d.free()
    finally:
        h.close()


def test_noalloc_first_use_under_capture():
    import ecalj_amd

    h = ecalj_amd.Handle(0, "f64")  # fresh handle: no scratch slot has ever been allocated
    try:
        n, lda, nb = 300, 300, 2
        say(f"zminv_noalloc_dev under graph capture, no warm-up: n={n} nbatch={nb}")
        rng = np.random.default_rng(7)
        a = _matrices(rng, n, lda, nb)
        src = h.to_dev(a)
        work = h.to_dev(np.zeros_like(a))
        h.sync()
        nbytes = a.nbytes
        h.graph_begin()
        h.dev_copy(work, src, nbytes)
        h.zminv_noalloc_dev(work, n, nbatch=nb, lda=lda)
        graph = h.graph_end()
        ref = None
        for rep in range(3):
            graph.launch()
            h.sync()
            out = work.download().reshape(a.shape, order="F")
            if ref is None:
                ref = out.copy()
                for b in range(nb):
                    inv = np.linalg.inv(a[:n, :, b])
                    assert np.abs(out[:n, :, b] - inv).max() / np.abs(inv).max() < 1e-9
            assert np.array_equal(out, ref)
            # eager work between replays, on the same scratch
            e1, e2 = _both(h, a[:100, :100, :1].copy(order="F"), 100, 100, 1)
            assert np.array_equal(e1, e2)
        graph.close()
        src.free()
        work.free()
    finally:
        h.close()
