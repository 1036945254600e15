"""Real-axis pole term of Sigma_c on the GPU (ecalj_sigc_real / _dev) against
the FP64 oracle (orc_sigc_real, pinned in test_sigc_real_oracle.py).
Bars: FP64 1e-10 relative; mixed forms (zmel stored as complex64, the term
itself in FP64) 1e-6."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(nw):
    aa, bb = 0.03, 1e-5 / 0.03
    fr = bb * (np.exp(aa * np.arange(nw + 2)) - 1.0)
    return np.concatenate([[0.0], 0.5 * (fr[:nw] + fr[1:nw + 1])])


def _case(seed, ngb, nm, ntq, nk, nw=270):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((ngb, nm, ntq, nk)) + 1j * rng.standard_normal((ngb, nm, ntq, nk))
    wr = rng.standard_normal((ngb, ngb, nw + 1, nk)) + 1j * rng.standard_normal((ngb, ngb, nw + 1, nk))
    ekc = np.sort(rng.uniform(-0.8, 0.6, (nm, nk)), axis=0)
    omega = rng.uniform(-0.7, 0.5, ntq)
    wk = rng.uniform(0.05, 0.2, nk)
    return z, wr, ekc, omega, wk, _mesh(nw)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("f32", 1e-6), ("bf16x3", 1e-6)])
@pytest.mark.parametrize("ngb,nm,ntq,nctot", [(40, 17, 6, 0), (97, 33, 9, 2), (130, 12, 3, 0)])
def test_sigc_real_host_vs_oracle(orc, prec, tol, ngb, nm, ntq, nctot):
    import ecalj_amd

    say("sigc_real", prec, ngb, nm, ntq)
    z, wr, ekc, omega, wk, fr = _case(ngb, ngb, nm, ntq, 1)
    ef, esmr = 0.01, 0.01
    ref, ne = orc.sigc_real(z[..., 0], wr[..., 0], fr, omega, ekc[:, 0], ef, esmr, wk[0], nctot=nctot)
    assert ne > 0
    h = ecalj_amd.Handle(0, prec)
    try:
        got = h.sigc_real(z[..., 0], wr[..., 0], fr, omega, ekc[:, 0], ef, esmr, wk[0], nctot=nctot)
    finally:
        h.close()
    assert np.abs(got - ref).max() / np.abs(ref).max() < tol


@pytest.mark.parametrize("shared", [False, True])
def test_sigc_real_batched_dev(orc, shared):
    import ecalj_amd

    ngb, nm, ntq, nk = 72, 21, 5, 4
    z, wr, ekc, omega, wk, fr = _case(7, ngb, nm, ntq, nk)
    if shared:
        wr = np.repeat(wr[..., :1], nk, axis=3)
    ef, esmr = -0.05, 0.01
    ref = np.zeros((ntq, ntq), complex, order="F")
    for k in range(nk):
        ref, _ = orc.sigc_real(z[..., k], wr[..., k], fr, omega, ekc[:, k], ef, esmr, wk[k], zsec=ref)
    h = ecalj_amd.Handle(0, "f64")
    try:
        dz = h.to_dev(z)
        dw = h.to_dev(wr[..., :1] if shared else wr)
        dsec = h.array((ntq, ntq))
        dsec.zero()
        h.sigc_real_dev(nk, dz, ngb, nm, ntq, dw, 0 if shared else ngb * ngb * len(fr), len(fr) - 1, fr, omega,
                        np.asfortranarray(ekc), ef, esmr, wk, dsec)
        got = dsec.download()
    finally:
        h.close()
    # per-k clipping (oracle) and one clip at the end (batched) agree: the diagonal is real
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10


def test_sigc_real_pole_outside_mesh():
    import ecalj_amd

    z, wr, ekc, omega, wk, _ = _case(3, 16, 8, 3, 1, nw=20)
    fr = _mesh(20)
    h = ecalj_amd.Handle(0, "f64")
    try:
        with pytest.raises(ecalj_amd.EcaljError) as e:
            h.sigc_real(z[..., 0], wr[..., 0], fr, omega, ekc[:, 0], 0.0, 0.01, 0.1)
        assert e.value.status == 1
    finally:
        h.close()


# bf16f8: the imaginary-axis part with fp8 cross terms (bar of tests/test_gpu_sigc_fp8.py)
@pytest.mark.parametrize("prec,tol", [("f64", 1e-10), ("bf16x3", 1e-4), ("bf16f8", 1e-3)])
def test_sigc_full_dev_vs_oracle(orc, prec, tol):
    """Imaginary + real axis in one czmelwc, k by k with the clip after each k,
    against orc_sigc_full called per k (the reference's order).  The real-axis
    W has complex diagonals, so the clip acts."""
    import ecalj_amd

    say("sigc_full", prec)
    ngb, nm, ntq, nk, niw = 64, 19, 6, 3, 6
    z, wr, ekc, omega, wk, fr = _case(11, ngb, nm, ntq, nk)
    rng = np.random.default_rng(12)
    a = rng.standard_normal((ngb, ngb, niw + 1, nk)) + 1j * rng.standard_normal((ngb, ngb, niw + 1, nk))
    wc = a + a.conj().transpose(1, 0, 2, 3)
    g = rng.standard_normal((niw + 1, nm, ntq, nk))
    ef, esmr = 0.0, 0.01
    ref = np.zeros((ntq, ntq), complex, order="F")
    for k in range(nk):
        ref, _ = orc.sigc_full(z[..., k], wc[..., k], g[..., k], wr[..., k], fr, omega, ekc[:, k], ef, esmr, wk[k],
                               zsec=ref)
    h = ecalj_amd.Handle(0, prec)
    try:
        cst = np.complex128 if prec == "f64" else np.complex64
        dz, dwc, dg, dwr = h.to_dev(z.astype(cst)), h.to_dev(wc), h.to_dev(g), h.to_dev(wr)
        dsec = h.array((ntq, ntq))
        dsec.zero()
        h.sigc_full_dev(nk, dz, ngb, nm, ntq, niw, dwc, dg, dwr, ngb * ngb * len(fr), len(fr) - 1, fr, omega,
                        np.asfortranarray(ekc), ef, esmr, wk, dsec)
        got = dsec.download()
    finally:
        h.close()
    assert np.abs(got - ref).max() / np.abs(ref).max() < tol
