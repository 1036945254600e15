"""synth.redraw_high_bands, the tool behind the measured bound on any learned
chi0 tail: the kept bands, the energies and the truncated chi0 are unchanged,
the eigenvectors stay orthonormal, and the full band sum changes."""
import dataclasses

import numpy as np

from ecalj_amd import synth


def test_redraw_keeps_the_truncated_sum(orc):
    from oracle import pipeline

    cfg = synth.CONFIGS["T3"].scaled(nu_chi0=12)
    qp = synth.make_qpoint(cfg)
    q2 = synth.redraw_high_bands(qp, cfg.nu_chi0, seed=3)
    keep = cfg.no + cfg.nu_chi0
    assert np.array_equal(q2.c_kq[:, :keep], qp.c_kq[:, :keep])
    assert not np.allclose(q2.c_kq[:, keep:], qp.c_kq[:, keep:])
    for k in range(cfg.nk):
        c = q2.c_kq[:, :, k]
        assert np.abs(c.conj().T @ c - np.eye(c.shape[1])).max() < 1e-12
    assert np.array_equal(q2.e_kq, qp.e_kq) and np.array_equal(q2.c_k, qp.c_k)
    # truncated chi0 identical; the full band sum is not
    chi_t1, _ = pipeline.chi0(qp)
    chi_t2, _ = pipeline.chi0(q2)
    assert np.array_equal(chi_t1, chi_t2)
    full = cfg.scaled(nu_chi0=0)
    chi_f1, _ = pipeline.chi0(dataclasses.replace(qp, cfg=full))
    chi_f2, _ = pipeline.chi0(dataclasses.replace(q2, cfg=full))
    assert np.abs(chi_f1 - chi_f2).max() > 1e-3 * np.abs(chi_f1).max()
