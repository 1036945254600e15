"""The Coulomb-matrix generalized eigensolve (zhgv, S/m_lapack.f90:64-104 as
called by hvccfp0, S/main_hvccfp0.f90:322-368) on the GPU against LAPACK
(scipy.linalg.eigh(a, b), which calls zhegvd) at the product-basis sizes of
the benchmark workloads (nI = 600 and 1200), a small odd size, exactly
degenerate eigenvalues and an indefinite a.

Bars: eigenvalues 1e-10 relative to the largest; b-orthonormality and the
residual a z - b z lambda at 1e-10; eigenvectors compared through the
b-orthogonal projector z_c z_c^H b of every cluster of equal eigenvalues
(unique whatever basis a solver picks inside a degenerate space)."""
import sys

import numpy as np
import pytest
import scipy.linalg

from ecalj_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(n, seed, degenerate=False, indefinite=False):
    rng = np.random.default_rng(seed)
    if degenerate:   # shells of exactly equal eigenvalues (q = 0 Coulomb shells)
        g = synth.gvectors(n + 1)[1:]
        lam = 4 * np.pi / (g ** 2).sum(1)
    else:            # 4 pi / |q + G|^2 of the synthetic Coulomb matrices (q = (0.1, 0.2, 0.3) 2 pi / a)
        g = synth.gvectors(n)
        q = np.array([0.1, 0.2, 0.3])
        lam = 4 * np.pi / ((((q + g) * 2 * np.pi / 10.68) ** 2).sum(1))
    if indefinite:
        lam = lam * np.where(np.arange(n) % 3 == 0, -1.0, 1.0)
    x = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    qm, _ = np.linalg.qr(x)
    v = (qm * lam) @ qm.conj().T
    y = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    o = np.eye(n) + 0.5 * (y @ y.conj().T) / n      # Hermitian positive definite overlap
    # only the upper triangles are read: scramble the strict lower ones
    vs, os_ = v.copy(order="F"), o.copy(order="F")
    il = np.tril_indices(n, -1)
    vs[il] = 7.0 + 3.0j
    os_[il] = -5.0 + 1.0j
    return v, o, vs, os_


def _check(v, o, evl, z, tol=1e-10):
    ref_e, ref_z = scipy.linalg.eigh(v, o)
    scale = np.abs(ref_e).max()
    e_err = np.abs(evl - ref_e).max() / scale
    orth = np.abs(z.conj().T @ o @ z - np.eye(len(evl))).max()
    res = np.abs(v @ z - (o @ z) * evl).max() / scale
    # projectors onto each cluster of equal eigenvalues
    perr = 0.0
    i = 0
    n = len(evl)
    while i < n:
        j = i + 1
        while j < n and abs(ref_e[j] - ref_e[i]) <= 1e-9 * scale:
            j += 1
        pg = z[:, i:j] @ z[:, i:j].conj().T @ o
        pr = ref_z[:, i:j] @ ref_z[:, i:j].conj().T @ o
        perr = max(perr, np.abs(pg - pr).max())
        i = j
    say(f"  n={n}: eig {e_err:.2e} orth {orth:.2e} resid {res:.2e} projector {perr:.2e}")
    assert e_err < tol and orth < tol and res < tol
    return perr


@pytest.mark.parametrize("n,degenerate,indefinite", [(157, False, False), (600, False, False),
                                                      (1200, False, False), (300, True, False),
                                                      (200, True, True)])
def test_zhgv_vs_lapack(n, degenerate, indefinite):
    import ecalj_amd

    say("zhgv", n, "degenerate" if degenerate else "", "indefinite" if indefinite else "")
    v, o, vs, os_ = _case(n, seed=n, degenerate=degenerate, indefinite=indefinite)
    h = ecalj_amd.Handle(0, "f64")
    try:
        evl, z = h.zhgv(vs, os_)
    finally:
        h.close()
    perr = _check(v, o, evl, z)
    # well-separated clusters: projectors to 1e-8 (their condition is 1/gap)
    assert perr < 1e-8


def test_zhgv_index_range_and_errors():
    import ecalj_amd

    n = 90
    v, o, vs, os_ = _case(n, seed=3)
    h = ecalj_amd.Handle(0, "f64")
    try:
        evl_all, z_all = h.zhgv(vs, os_)
        evl, z = h.zhgv(vs, os_, il=11, iu=40)
        assert evl.shape == (30,) and z.shape == (n, 30)
        assert np.allclose(evl, evl_all[10:40], rtol=0, atol=1e-12 * np.abs(evl_all).max())
        # not positive definite b: the reference's zhegvx reports info > n
        bad = os_.copy(order="F")
        bad[5, 5] = -1.0
        with pytest.raises(ecalj_amd.EcaljError) as e:
            h.zhgv(vs, bad)
        assert e.value.status == 7
        evl2, _ = h.zhgv(vs, os_)     # the handle works on after the error
        assert np.array_equal(evl2, evl_all)
    finally:
        h.close()


def test_coulomb_from_matrix_feeds_the_w_build(orc):
    """hvccfp0's use: eigenvalues of vcoul in descending order from zhgv(-vcoul, oo),
    then Readvcoud; the W build from them equals the oracle's from LAPACK's."""
    import ecalj_amd

    n = 200
    v, o, vs, os_ = _case(n, seed=11)
    h = ecalj_amd.Handle(0, "f64")
    try:
        vcoud, zc = h.coulomb_from_matrix(vs, os_)
        assert np.all(np.diff(vcoud) <= 0)
        ref_e, ref_z = scipy.linalg.eigh(-v, o)
        assert np.abs(vcoud + ref_e).max() < 1e-10 * np.abs(ref_e).max()
        rng = np.random.default_rng(2)
        x = rng.standard_normal((n, 30)) + 1j * rng.standard_normal((n, 30))
        chi = np.asfortranarray(np.stack([-(x @ x.conj().T) * 1e-3 / (1 + i) for i in range(3)], axis=2))
        w = h.wc_build(chi, nspin=1)
        # the oracle from LAPACK's eigenpairs: W_c depends only on v(q) = Z diag(vcoud) Z^H,
        # so the different eigenvector bases inside degenerate spaces do not matter
        w_ref = orc.wc_build(chi, np.sqrt(-ref_e), ref_z, nspin=1)
        err = np.abs(w - w_ref).max() / np.abs(w_ref).max()
        say(f"  W from zhgv's v(q): rel {err:.2e}")
        assert err < 1e-10
    finally:
        h.close()
