"""One q-point of the QSGW chi0 -> W -> Sigma_c path on the GPU.

Mirrors the reference drivers with device-resident data:
  hrcxq (S/main_hrcxq.f90:101-120) -> x0kf_zxq (S/x0kf_v4h.f90:144-328):
      zmel for all k at once (conj), x0gemm over all k at once, dpsion_chiq;
  WVIllwI (S/m_llw.f90:265-449) for q /= 0;
  sxcf_scz_correlation (S/m_sxcf_gemm.f90:255-657): zmel for all internal k,
      imaginary-axis weights, frequency contraction, final product;
  hqpe_sc (S/main_hqpe.sc.f90:160-161).
The frequency list for W and Sigma_c is {0, nu_1..nu_nf}.  W_c at the
internal k-points is D_k W_c(q) D_k^H with seeded unit phases D_k
(``synth.wc_phases``), standing in for the W files of the other q-points.
Synthetic inputs and index lists (x0kf_v4hz_init) are made in the
constructor; ``step`` issues device work plus the per-bin gather lists that
x0gemm builds on the host.
"""
from __future__ import annotations

import math

import numpy as np

from . import Handle, qp_energy, synth

RY = synth.RYDBERG_EV


def gauss01(n):
    """Gauss-Legendre nodes/weights on (0,1), ascending (gauss, S/mate.f90)."""
    t, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (t + 1.0), 0.5 * w


class QPointGPU:
    """tail_model: a learned_tail.TailModel; chi0 then sums the unoccupied bands
    up to the model's nu_t and adds the predicted tail on the device.

    overlap: Sigma_c's zmel and weights (independent of chi0 and W) run on a
    second handle's stream while W_c(q) is built; the two streams join before
    the contraction (``Handle.stream_wait``).  Same results, same kernels.

    real_axis: also build chi(omega) and W_c(omega) on the real mesh freq_r
    (dpsion_chiq real branch, WVRllwR) from the same rcxq, and add the
    real-axis pole term to Sigma_c (sxcf_scz_correlation, S/m_sxcf_gemm.f90:
    501-594) with the imaginary-axis part in one contraction.  W_c(omega) of
    q serves every internal k (the pole term reads a few frequencies each)."""

    def __init__(self, qp: synth.QPoint, precision: str = "f64", device: int = 0, tail_model=None,
                 real_axis: bool = False, overlap: bool = False):
        self.qp = qp
        self.cfg = cfg = qp.cfg
        self.h = h = Handle(device, precision)
        c128, f64 = np.complex128, np.float64
        # bulk intermediates (zmel, rcxq) in the contraction's storage type
        cst = c128 if precision == "f64" else np.complex64
        self.cst = cst
        no, nbt, nb, ni, nk, ntq = cfg.no, cfg.nbt, cfg.nb, cfg.ni, cfg.nk, cfg.nq
        # unoccupied bands in the chi0 band sum (nband_chi0, S/m_readgwinput.f90:34-35)
        self.nuc = nu = cfg.nu_chi0 if cfg.nu_chi0 else cfg.nu
        if tail_model is not None:
            self.nuc = nu = tail_model.nu_t
        self.nw1 = cfg.nf + 1
        # module state
        h.set_ppb(cfg.nv_list, cfg.mdim_list, qp.ppb)
        h.set_coulomb(qp.vcoud, qp.zcousq)
        # eigenvectors on the device, one block per k
        self.c_occ = h.to_dev(qp.c_k[:, :no, :], c128)       # nb x no x nk
        self.c_unocc = h.to_dev(qp.c_kq[:, no:no + nu, :], c128)    # nb x nu x nk
        self.c_qmk = h.to_dev(qp.c_qmk, c128)                # nb x nbt x nk
        self.c_q = h.to_dev(qp.c_q[:, :ntq], c128)           # nb x ntq
        # frequencies
        self.freqx, self.wx = gauss01(cfg.nf)
        self.freqw = (1.0 - self.freqx) / self.freqx
        self.expa = np.exp(-cfg.ua ** 2 * self.freqw ** 2)
        self.nu = np.concatenate([[0.0], self.freqw])
        # histogram and x0 index lists
        self.fr = synth.frhis(qp)
        self.nwhis = len(self.fr) - 1
        self.ix = synth.x0kf_index(synth.pair_weights(qp, self.fr, nu))
        # buffers
        self.zmel_chi = h.array((ni, no, nu, nk), cst)
        self.rcxq = h.array((ni, ni, self.nwhis + 1), cst)
        self.chi = h.array((ni, ni, self.nw1), c128)
        self.wc = h.array((ni, ni, self.nw1), c128)
        self.wk = h.array((ni, ni, self.nw1, nk), c128)
        self.zmel_sig = h.array((ni, nbt, ntq, nk), cst)
        self.wgt = h.array((self.nw1, nbt, ntq, nk), f64)
        self.zsec = h.array((ntq, ntq), c128)
        # Sigma weights inputs
        self.d_omega = h.to_dev(np.tile(qp.e_q[:ntq] / RY, nk), f64)
        self.d_ekc = h.to_dev((qp.e_qmk / RY), f64)         # nbt x nk
        self.d_wk = h.to_dev(qp.wk, f64)
        self.d_x = h.to_dev(self.freqx, f64)
        self.d_wx = h.to_dev(self.wx, f64)
        self.d_expa = h.to_dev(self.expa, f64)
        self.phase = h.to_dev(synth.wc_phases(cfg), c128)   # ni x nk
        self.real_axis = real_axis
        if real_axis:
            if tail_model is not None:
                raise ValueError("the learned tail models the imaginary axis only")
            self.freqr = synth.real_mesh(qp, self.fr)
            self.nwr = len(self.freqr) - 1
            self.chi_r = h.array((ni, ni, self.nwr + 1), c128)
            self.w_r = h.array((ni, ni, self.nwr + 1), c128)
            self.d_omega1 = qp.e_q[:ntq] / RY
            self.ekc_h = np.asfortranarray(qp.e_qmk / RY)
        self.h2 = None
        if overlap:
            self.h2 = Handle(device, precision)
            self.h2.set_ppb(cfg.nv_list, cfg.mdim_list, qp.ppb)
        self.tail = None
        if tail_model is not None:
            from .learned_tail import LearnedTail
            self.tail = LearnedTail(h, tail_model)

    # ---------------------------------------------------------------- chi0
    def chi0(self):
        cfg = self.cfg
        no, nu, nb, nk = cfg.no, self.nuc, cfg.nb, cfg.nk
        # zmel(I, n_occ, m_unocc) for every k, conj (zmelconjg=.true., S/x0kf_v4h.f90:273-275)
        self.h.zmel_dev(self.c_occ, self.c_unocc, self.zmel_chi, nm=no, nq=nu, ndima=nb, nbatch=nk,
                        is_m_basis=True, zmelconjg=True)
        self.chi0_pairs(self.ix, self.chi)
        if self.tail is not None:
            self.tail.apply(self.chi, cfg.ni, self.chi)

    def chi0_pairs(self, ix, out):
        """x0gemm over the pairs of the index lists ix (from the zmel of chi0()) + Hilbert -> out."""
        cfg = self.cfg
        no, nu, ni, nk = cfg.no, self.nuc, cfg.ni, cfg.nk
        self.rcxq.zero()
        self.h.x0gemm_batched_dev(self.rcxq, ni, self.nwhis, nk, ix.nkmin, ix.nkmax, self.zmel_chi, ni * no * nu,
                                  ni, nu, ix)
        self.h.dpsion_imag_dev(self.rcxq, ni, ni, self.nwhis, self.fr, self.nu, out)
        if self.real_axis:
            self.h.dpsion_real_dev(self.rcxq, ni, ni, self.nwhis, self.fr, self.freqr, self.chi_r)

    def wbuild(self):
        cfg = self.cfg
        self.h.wc_build_dev(self.chi, self.wc, cfg.ni, self.nw1, nspin=cfg.nspin)
        if self.real_axis:
            self.h.wc_build_real_dev(self.chi_r, self.w_r, cfg.ni, self.nwr + 1, nspin=cfg.nspin)

    def make_wk(self):
        """W_c at the internal k: D_k W_c(q) D_k^H (inputs for Sigma_c)."""
        cfg = self.cfg
        self.h.phase_similarity(self.wc, cfg.ni, self.nw1, self.phase, cfg.nk, self.wk)

    # ---------------------------------------------------------------- Sigma_c
    def sigc_prep(self, h=None):
        """Sigma_c inputs that do not depend on W: zmel of the internal k and the weights."""
        h = self.h if h is None else h
        cfg = self.cfg
        nb, nbt, ntq, nk = cfg.nb, cfg.nbt, cfg.nq, cfg.nk
        # zmel(I, n_middle, m_end): middle states at q-k, end states at q shared (S/m_sxcf_gemm.f90:395-399)
        h.zmel_dev(self.c_qmk, self.c_q, self.zmel_sig, nm=nbt, nq=ntq, ndima=nb, nbatch=nk,
                   is_m_basis=True, zmelconjg=False, cphiq_stride=0)
        h.sigc_weights_dev(nk, cfg.nf, nbt, ntq, self.d_omega, self.d_ekc, cfg.esmr, cfg.ua, self.d_wk,
                           self.d_x, self.d_wx, self.d_expa, self.wgt)

    def sigc(self):
        self.sigc_prep()
        self.sigc_contract()

    def sigc_contract(self):
        cfg = self.cfg
        ni, nbt, ntq, nk = cfg.ni, cfg.nbt, cfg.nq, cfg.nk
        self.zsec.zero()
        if self.real_axis:
            self.h.sigc_full_dev(nk, self.zmel_sig, ni, nbt, ntq, cfg.nf, self.wk, self.wgt, self.w_r, 0, self.nwr,
                                 self.freqr, self.d_omega1, self.ekc_h, 0.0, cfg.esmr, self.qp.wk,
                                 self.zsec)
        else:
            self.h.sigc_dev(nk, self.zmel_sig, ni, nbt, ntq, cfg.nf, self.wk, self.wgt, self.zsec)

    def step(self):
        """chi0(q) + W_c(q) + Sigma_c(q): the per-q-point unit of work."""
        self.chi0()
        if self.h2 is None:
            self.wbuild()
            self.sigc()
            return
        # zmel/weights of Sigma_c beside the W build; the event after chi0 also
        # orders them after the previous step's contraction, which read them
        self.h2.stream_wait(self.h)
        self.sigc_prep(self.h2)
        self.wbuild()
        self.h.stream_wait(self.h2)
        self.sigc_contract()

    def capture_step(self):
        """A HIP graph of step() on one stream (call after one eager step of this
        q-point).  The host-side index lists are fixed at capture; a replay
        recomputes the q-point from the device buffers, e.g. in the next QSGW
        iteration."""
        if self.h2 is not None:
            self.h.stream_wait(self.h2)
        self.h.graph_begin()
        try:
            self.chi0()
            self.wbuild()
            self.sigc()
        finally:
            g = self.h.graph_end()
        return g

    def close(self):
        if self.h2 is not None:
            self.h2.sync()
            self.h2.close()
            self.h2 = None
        self.h.close()

    # ---------------------------------------------------------------- results
    def results(self):
        self.h.sync()
        ntq = self.cfg.nq
        chi = self.chi.download()
        wc = self.wc.download()
        zsec = self.zsec.download()
        sex, sexcore, vxc = synth.sigma_inputs(self.qp)
        eqp = qp_energy(self.qp.e_q[:ntq], sex, sexcore, zsec, vxc)
        return chi, wc, zsec, eqp


def flops_per_q(cfg, nwhis_pairs_bins=None):
    """Algorithmic FLOPs of one q-point as executed (complex MAC = 8, real x complex = 4)."""
    nk, nb, no, ni, nw1, ntq, nbt = cfg.nk, cfg.nb, cfg.no, cfg.ni, cfg.nf + 1, cfg.nq, cfg.nbt
    nu = cfg.nu_chi0 if cfg.nu_chi0 else cfg.nu
    f = {}
    na = cfg.natom  # atom blocks: sum_a nv_a^2 mdim_a = nb^2 ni / na^2, sum_a mdim_a nv_a = ni nb / na
    f["zmel_chi"] = nk * (4.0 * nb * nb * ni * no / na ** 2 + 8.0 * ni * no * nu * nb / na)
    f["x0"] = 8.0 * ni * ni * (nwhis_pairs_bins if nwhis_pairs_bins is not None else nk * no * nu)
    f["hilbert"] = 0.0  # memory-bound, counted in bytes
    f["wbuild"] = nw1 * (8.0 * ni ** 3 + 4 * 8.0 * ni ** 3)
    f["zmel_sig"] = 4.0 * nb * nb * ni * ntq / na ** 2 + nk * 8.0 * ni * nbt * ntq * nb / na
    f["sigc"] = nk * 8.0 * ni * ni * nw1 * nbt * ntq
    f["sigc_final"] = nk * 8.0 * ntq * ntq * ni * nbt
    return f
