"""Parity at the real shapes of the benchmark workloads.

Every shape parameter of the workload is kept -- nI, nb, the band windows,
nf and ntq -- and only the number of k-points is cut, so that the FP64
oracle (OpenBLAS GEMMs, as the reference's MKL ones) finishes in tens of
seconds on the host:

* C2  (nI = 600, nb = 180, 9 + 150 bands, nf = 16): 4 of 64 k-points;
* C3  (nI = 1200, nb = 320, 18 + 300 bands, nf = 20): 2 of 64 k-points;
* C4  (one q-point at nI = 1600, nb = 432, per-k windows (24 + 400) and
  (13 + 187), ntq = 16, nf = 20, grouped Sigma_c kernel): 2 of 27 k-points;
* C5  (nI = 1200, nb = 832 in 8 atoms, 8 + 800 bands, chi0 truncated at 200,
  nf = 20, ntq = 16, learned tail applied): 2 of 8 k-points.

Bars: the FP64 form 1e-10 relative on chi0, W_c and Sigma_c and 1e-8 eV on
QP energies; the mixed forms within 1 meV of the oracle's QP energies, with
per-result relative bars (f32 2e-5 and bf16x3 1e-4 on chi0; W_c and Sigma_c
within 10x that; bf16f8 1e-3 on Sigma_c)."""
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu

TOL = {"f32": 2e-5, "bf16x3": 1e-4, "bf16f8": 1e-4}
SIG_TOL = {"f32": 2e-4, "bf16x3": 1e-3, "bf16f8": 1e-3}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _oracle(qp):
    from oracle import pipeline

    say(f"oracle: {qp.cfg.name} nk={qp.cfg.nk} nI={qp.cfg.ni} ...")
    ref = pipeline.run(qp)
    say(f"oracle: {qp.cfg.name} done")
    return ref


def _run_uniform(qp, prec, tail_model=None, mode=0):
    from ecalj_amd.qsgw import QPointGPU

    say(f"gpu: {qp.cfg.name} {prec} tuner mode {mode}")
    g = QPointGPU(qp, prec, tail_model=tail_model)
    try:
        g.h.set_tuning(mode)
        g.chi0()
        g.wbuild()
        g.make_wk()
        g.sigc()
        return g.results()
    finally:
        g.h.close()


def _check_f64(res, ref):
    chi, wc, zsec, eqp = res
    errs = rel(chi, ref.chi), rel(wc, ref.wc), rel(zsec, ref.zsec), float(np.abs(eqp - ref.eqp).max())
    say(f"  f64: chi {errs[0]:.2e} W {errs[1]:.2e} Sigma {errs[2]:.2e} |dEqp| {errs[3]:.2e} eV")
    assert errs[0] < 1e-10 and errs[1] < 1e-10 and errs[2] < 1e-10
    assert errs[3] < 1e-8


def _check_mixed(res, ref, prec):
    chi, wc, zsec, eqp = res
    mev = float(np.abs(eqp - ref.eqp).max() * 1e3)
    errs = rel(chi, ref.chi), rel(wc, ref.wc), rel(zsec, ref.zsec)
    say(f"  {prec}: chi {errs[0]:.2e} W {errs[1]:.2e} Sigma {errs[2]:.2e} |dEqp| {mev:.4f} meV")
    assert mev < 1.0
    assert errs[0] < TOL[prec]
    assert errs[1] < 10 * TOL[prec]
    assert errs[2] < SIG_TOL[prec]


# ------------------------------------------------------------------ C2
@pytest.fixture(scope="module")
def c2():
    qp = synth.make_qpoint(synth.CONFIGS["C2"].scaled(nk=4))
    return qp, _oracle(qp)


def test_c2_f64(c2):
    qp, ref = c2
    _check_f64(_run_uniform(qp, "f64"), ref)


@pytest.mark.parametrize("prec,mode", [("bf16x3", 0), ("bf16x3", 4), ("bf16x3", 6), ("bf16x3", 8), ("bf16x3", 10),
                                       ("bf16x3", 12), ("f32", 0)])
def test_c2_mixed(c2, prec, mode):
    """Mixed forms; bf16x3 with the tuner's choice and with the frequency-outer
    (4), three-product (6), four-product frequency-inner (8), three-product
    8-wave (10) and three-product 32x32x16 (12) Sigma_c kernels forced."""
    qp, ref = c2
    _check_mixed(_run_uniform(qp, prec, mode=mode), ref, prec)


# ------------------------------------------------------------------ C3
@pytest.fixture(scope="module")
def c3():
    qp = synth.make_qpoint(synth.CONFIGS["C3"].scaled(nk=2))
    return qp, _oracle(qp)


def test_c3_f64(c3):
    qp, ref = c3
    _check_f64(_run_uniform(qp, "f64"), ref)


@pytest.mark.parametrize("prec,mode", [("bf16x3", 0), ("bf16x3", 4), ("bf16x3", 6), ("bf16x3", 8), ("bf16x3", 12),
                                       ("f32", 0), ("bf16f8", 0)])
def test_c3_mixed(c3, prec, mode):
    qp, ref = c3
    _check_mixed(_run_uniform(qp, prec, mode=mode), ref, prec)


# ------------------------------------------------------------------ C4
@pytest.fixture(scope="module")
def c4():
    cfg = synth.CONFIGS["C4"].scaled(nk=2, ni=1600, seed=20261206 + 77)
    qp = synth.make_uneven_qpoint(cfg, no_range=(8, 24), nu_range=(150, 400), windows=([24, 13], [400, 187]))
    return qp, _oracle(qp)


def _run_uneven(qp, prec, mode=0):
    from ecalj_amd.batched import UnevenQPointGPU

    say(f"gpu: C4 uneven {prec} tuner mode {mode}")
    g = UnevenQPointGPU(qp, prec)
    try:
        g.h.set_tuning(mode)
        g.chi0()
        g.wbuild()
        g.make_wk()
        g.sigc()
        return g.results()
    finally:
        g.h.close()


def test_c4_f64(c4):
    qp, ref = c4
    _check_f64(_run_uneven(qp, "f64"), ref)


@pytest.mark.parametrize("prec,mode", [("bf16x3", 0), ("bf16x3", 1), ("bf16x3", 4), ("bf16x3", 5), ("bf16f8", 0)])
def test_c4_grouped_mixed(c4, prec, mode):
    """The grouped persistent Sigma_c kernel over both k-points' tiles (the
    tuner's choice, the 128x64 variant and the three-product form in both
    8-wave layouts forced), at nI = 1600 and a 424-band window."""
    qp, ref = c4
    _check_mixed(_run_uneven(qp, prec, mode), ref, prec)


# ------------------------------------------------------------------ C5
def _tail_model(cfg, seed=5):
    """A seeded, untrained tail model of the C5 shape: parity of features,
    bf16 inference and reconstruction does not depend on training."""
    from ecalj_amd import learned_tail as lt
    from ecalj_amd import mlp_nparam

    nw = cfg.nf + 1
    dims = lt.tail_dims(nw)
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in zip(dims[:-1], dims[1:]):
        parts.append((rng.standard_normal((b, a)) / np.sqrt(a)).ravel())   # W (row-major, out x in)
        parts.append(0.1 * rng.standard_normal(b))                          # b
    params = np.concatenate(parts).astype(np.float32)
    assert params.size == mlp_nparam(dims)
    return lt.TailModel(dims, params, np.zeros(nw, np.float32), np.ones(nw, np.float32),
                        np.full(nw, 0.5, np.float32), np.full(nw, 0.2, np.float32), iref=nw - 1,
                        nu_t=cfg.nu_chi0)


@pytest.fixture(scope="module")
def c5():
    cfg = synth.CONFIGS["C5"].scaled(nk=2)
    qp = synth.make_qpoint(cfg)                       # chi0 truncated at nu_chi0 = 200 in the oracle too
    return qp, _oracle(qp)


def test_c5_truncated_f64(c5):
    """The truncated band sum (200 of 800 unoccupied bands) and what follows from it."""
    qp, ref = c5
    _check_f64(_run_uniform(qp, "f64"), ref)


def test_c5_full_band_bf16x3(c5):
    """The full band sum (the judge and fallback of the learned form) in the mixed form."""
    import dataclasses

    from oracle import pipeline

    qp, _ = c5
    qf = dataclasses.replace(qp, cfg=qp.cfg.scaled(nu_chi0=0))
    chi_ref, _ = pipeline.chi0(qf)
    chi = _run_uniform(qf, "bf16x3")[0]
    say(f"  full-band chi0 bf16x3 rel {rel(chi, chi_ref):.2e}")
    assert rel(chi, chi_ref) < TOL["bf16x3"]


def test_c5_learned_tail_applied(c5):
    """chi_L = chi_T + tail at the C5 shape: the GPU's features, bf16 MLP
    inference and reconstruction against a host restatement from the
    oracle's chi_T (features d_j/d_0 and log d_0; the FP64 oracle MLP; tail
    diagonal t = max(0, d_0 (mu_y + sd_y y)); off-diagonal
    chi_T(I,J,iref) sqrt(t_I t_J) / sqrt(d_iref(I) d_iref(J))); and the QP
    energies of the learned form within 1 meV of those from the host chi_L."""
    import oracle as orc
    from ecalj_amd.learned_tail import qp_from_chi
    from ecalj_amd.qsgw import QPointGPU

    qp, ref = c5
    cfg = qp.cfg
    m = _tail_model(cfg)
    ni, nw = cfg.ni, cfg.nf + 1
    g = QPointGPU(qp, "f64", tail_model=m)
    try:
        g.chi0()
        chi_l_gpu = g.chi.download()
        ct = ref.chi
        d = -np.real(np.einsum("iij->ij", ct))
        assert (d[:, 0] > 0).all()
        x = np.concatenate([d[:, 1:] / d[:, :1], np.log(d[:, :1])], axis=1)
        x = (x - m.mu_x.astype(np.float64)) / m.sd_x.astype(np.float64)
        y = orc.mlp_forward(m.dims, m.params.astype(np.float64), x)
        t = np.maximum(0.0, d[:, :1] * (m.mu_y.astype(np.float64) + m.sd_y.astype(np.float64) * y))
        dref = d[:, m.iref]
        rho = ct[:, :, m.iref] / np.sqrt(dref[:, None] * dref[None, :])
        chi_l = ct + rho[:, :, None] * np.sqrt(t[:, None, :] * t[None, :, :])
        tail_err = rel(chi_l_gpu - ct, chi_l - ct)
        say(f"  learned chi_L rel {rel(chi_l_gpu, chi_l):.2e}, tail part rel {tail_err:.2e}")
        assert rel(chi_l_gpu, chi_l) < 1e-4
        assert tail_err < 1e-3      # bf16 hi/lo inference of the MLP
        e_gpu = qp_from_chi(g, g.chi)
        host = g.h.to_dev(np.asfortranarray(chi_l))
        e_host = qp_from_chi(g, host)
        mev = float(np.abs(e_gpu - e_host).max() * 1e3)
        say(f"  learned form: |dEqp| against the host chi_L {mev:.4f} meV")
        assert mev < 1.0
    finally:
        g.h.close()
