"""The hot-path entry points called from Fortran through ecalj_gpu_mod.f90
(tests/fortran/hotpath_call.f90, built by build()): per k ecalj_zmel +
ecalj_x0gemm, ecalj_dpsion_imag, ecalj_wc_build, then per internal k
ecalj_zmel + ecalj_sigc_weights + ecalj_sigc, as a driver (x0kf_zxq,
WVIllwI, get_correlation_block) makes those calls, against the FP64 oracle
on the same seeded q-point.  Bars: FP64 1e-10 relative; bf16x3 1e-4."""
import os
import subprocess
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "hotpath_call")


def _write_inputs(path, qp, prec):
    cfg = qp.cfg
    fr = synth.frhis(qp)
    nwhis = len(fr) - 1
    ix = synth.x0kf_index(synth.pair_weights(qp, fr))
    x, w = np.polynomial.legendre.leggauss(cfg.nf)
    freqx, wx = 0.5 * (x + 1), 0.5 * w
    freqw = (1 - freqx) / freqx
    expa = np.exp(-cfg.ua ** 2 * freqw ** 2)
    freqi = np.concatenate([[0.0], freqw])
    no, nu, nbt, ntq, nk = cfg.no, cfg.nu, cfg.nbt, cfg.nq, cfg.nk
    ints = np.array([prec, cfg.natom, cfg.nb, cfg.ni, nk, no, nu, ntq, nbt, nwhis, cfg.nf + 1, len(ix.itc),
                     len(ix.whwc), qp.ppb.size], np.int32)
    c = lambda a: np.asfortranarray(a, np.complex128).ravel(order="F")
    d = lambda a: np.asfortranarray(a, np.float64).ravel(order="F")
    i = lambda a: np.asarray(a, np.int32).ravel()
    with open(path, "wb") as f:
        for a in (ints, i(cfg.nv_list), i(cfg.mdim_list), d(qp.ppb), d(qp.vcoud), c(qp.zcousq),
                  c(np.ones(cfg.natom)), c(qp.c_k[:, :no, :]), c(qp.c_kq[:, no:no + nu, :]), i(ix.icounkmin),
                  i(ix.icounkmax), i(ix.iwini), i(ix.iwend), i(ix.itc), i(ix.itpc), i(ix.jpmc), i(ix.icouini),
                  d(ix.whwc), d(fr), d(freqi), c(qp.c_qmk), c(qp.c_q[:, :ntq]), d(qp.e_q[:ntq] / synth.RYDBERG_EV),
                  d(qp.e_qmk / synth.RYDBERG_EV), d(qp.wk), d(freqx), d(wx), d(expa),
                  np.array([cfg.esmr, cfg.ua], np.float64)):
            a.tofile(f)
    return freqx, wx, expa


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (2, 1e-4)])
def test_fortran_drives_the_hot_path(orc, tmp_path, prec, tol):
    from oracle import pipeline

    print(f"fortran hot path: T3 q-point, precision {prec}", flush=True)
    sys.stdout.flush()
    qp = synth.make_qpoint(synth.CONFIGS["T3"])
    cfg = qp.cfg
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    freqx, wx, expa = _write_inputs(fin, qp, prec)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    ni, nw1, ntq = cfg.ni, cfg.nf + 1, cfg.nq
    out = np.fromfile(fout, np.complex128)
    chi = out[:ni * ni * nw1].reshape((ni, ni, nw1), order="F")
    wc = out[ni * ni * nw1:2 * ni * ni * nw1].reshape((ni, ni, nw1), order="F")
    zsec = out[2 * ni * ni * nw1:].reshape((ntq, ntq), order="F")
    # the oracle, composed as the Fortran calls are (one W_c for every internal k)
    chi_ref, _ = pipeline.chi0(qp)
    wc_ref = orc.wc_build(chi_ref, qp.vcousq, qp.zcousq, nspin=cfg.nspin)
    zsec_ref = np.zeros((ntq, ntq), complex, order="F")
    for k in range(cfg.nk):
        z = orc.zmel(cfg.nv_list, cfg.mdim_list, [1.0] * cfg.natom, qp.ppb, qp.c_qmk[:, :, k], qp.c_q[:, :ntq])
        g = orc.wgtim(qp.e_q[:ntq] / synth.RYDBERG_EV, qp.e_qmk[:, k] / synth.RYDBERG_EV, cfg.esmr, qp.wk[k],
                      freqx, wx, expa, ua=cfg.ua)
        zsec_ref = orc.sigc_imag(z, wc_ref, g, zsec_ref)
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    print(f"  chi {rel(chi, chi_ref):.2e} W {rel(wc, wc_ref):.2e} Sigma {rel(zsec, zsec_ref):.2e}", flush=True)
    assert rel(chi, chi_ref) < tol
    assert rel(wc, wc_ref) < tol
    assert rel(zsec, zsec_ref) < 10 * tol
