"""q-points whose k-points carry their own band windows (the batched uneven workload).

ecalj's chi0 and Sigma_c loops run over k-points whose band ranges differ:
x0kf_v4hz_init records per k the occupied range nkmin..nkmax and the
unoccupied range nkqmin..nkqmax (S/x0kf_v4h.f90:91-137), and sxcf_scz_count
gives every internal k its own middle-state window (nstti/nstte,
S/sxcf_fal2_count.sc.f90:118-236).  Here:

* zmel is formed per k at its own shape (one launch per k, so no padding),
  into a per-k slot for chi0 and packed one after the other for Sigma_c;
* x0gemm runs once over all k, with the per-k occupied ranges (its bins
  already group pairs of uneven count);
* Sigma_c runs as one grouped persistent kernel over the (k, row tile,
  column tile) list of every k (ecalj_sigc_grouped_dev), tile shape from the
  library's tuner;
* the whole step can be captured in a HIP graph and replayed (the QSGW
  iterations repeat every q-point at unchanged shapes).
"""
from __future__ import annotations

import numpy as np

from . import Handle, qp_energy, synth
from .qsgw import gauss01

RY = synth.RYDBERG_EV


class UnevenQPointGPU:
    def __init__(self, qp: synth.QPoint, precision: str = "f64", device: int = 0):
        self.qp = qp
        self.cfg = cfg = qp.cfg
        self.h = h = Handle(device, precision)
        c128, f64 = np.complex128, np.float64
        cst = c128 if precision == "f64" else np.complex64
        self.cst = cst
        self.es = np.dtype(cst).itemsize
        nb, ni, nk, ntq, nbt = cfg.nb, cfg.ni, cfg.nk, cfg.nq, cfg.nbt
        self.no_k, self.nu_k = synth.band_windows(qp)
        self.nbt_k = self.no_k + self.nu_k
        self.nw1 = cfg.nf + 1
        h.set_ppb(cfg.nv_list, cfg.mdim_list, qp.ppb)
        h.set_coulomb(qp.vcoud, qp.zcousq)
        self.c_k = h.to_dev(qp.c_k, c128)       # nb x nbt x nk
        self.c_kq = h.to_dev(qp.c_kq, c128)
        self.c_qmk = h.to_dev(qp.c_qmk, c128)
        self.c_q = h.to_dev(qp.c_q[:, :ntq], c128)
        self.freqx, self.wx = gauss01(cfg.nf)
        self.freqw = (1.0 - self.freqx) / self.freqx
        self.expa = np.exp(-cfg.ua ** 2 * self.freqw ** 2)
        self.nu = np.concatenate([[0.0], self.freqw])
        self.fr = synth.frhis(qp)
        self.nwhis = len(self.fr) - 1
        self.ix = synth.x0kf_index(synth.pair_weights(qp, self.fr))
        # chi0: per-k zmel slots of the largest window
        self.zstride = ni * int((self.no_k * self.nu_k).max())
        self.nqtot = int(self.nu_k.max())
        self.zmel_chi = h.array(self.zstride * nk, cst)
        self.rcxq = h.array((ni, ni, self.nwhis + 1), cst)
        self.chi = h.array((ni, ni, self.nw1), c128)
        self.wc = h.array((ni, ni, self.nw1), c128)
        self.wk = h.array((ni, ni, self.nw1, nk), c128)
        # Sigma_c: zmel and weights packed per k
        self.zoff = np.concatenate([[0], np.cumsum(ni * self.nbt_k * ntq)])
        self.goff = np.concatenate([[0], np.cumsum(self.nw1 * self.nbt_k * ntq)])
        self.eoff = np.concatenate([[0], np.cumsum(self.nbt_k)])
        self.zmel_sig = h.array(int(self.zoff[-1]), cst)
        self.wgt = h.array(int(self.goff[-1]), f64)
        self.zsec = h.array((ntq, ntq), c128)
        self.d_omega = h.to_dev(qp.e_q[:ntq] / RY, f64)
        self.d_ekc = h.to_dev(np.concatenate([qp.e_qmk[:self.nbt_k[k], k] for k in range(nk)]) / RY, f64)
        self.d_wk = h.to_dev(qp.wk, f64)
        self.d_x = h.to_dev(self.freqx, f64)
        self.d_wx = h.to_dev(self.wx, f64)
        self.d_expa = h.to_dev(self.expa, f64)
        self.phase = h.to_dev(synth.wc_phases(cfg), c128)

    def _col(self, arr, k, band):
        """Device address of column `band` (0-based) of eigenvector block k (nb x nbt x nk)."""
        cfg = self.cfg
        return arr.ptr + 16 * (cfg.nb * (cfg.nbt * k + band))

    def chi0(self):
        cfg = self.cfg
        ni, nb, nk = cfg.ni, cfg.nb, cfg.nk
        for k in range(nk):  # zmel(I, n <= no_k, m in the k's unoccupied window), conj (S/x0kf_v4h.f90:273-275)
            no, nu = int(self.no_k[k]), int(self.nu_k[k])
            self.h.zmel_dev(self._col(self.c_k, k, 0), self._col(self.c_kq, k, no),
                            self.zmel_chi.ptr + self.es * self.zstride * k, nm=no, nq=nu, ndima=nb, nbatch=1,
                            is_m_basis=True, zmelconjg=True)
        self.rcxq.zero()
        ix = self.ix
        self.h.x0gemm_batched_dev(self.rcxq, ni, self.nwhis, nk, ix.nkmin, ix.nkmax, self.zmel_chi, self.zstride,
                                  ni, self.nqtot, ix)
        self.h.dpsion_imag_dev(self.rcxq, ni, ni, self.nwhis, self.fr, self.nu, self.chi)

    def wbuild(self):
        cfg = self.cfg
        self.h.wc_build_dev(self.chi, self.wc, cfg.ni, self.nw1, nspin=cfg.nspin)

    def make_wk(self):
        cfg = self.cfg
        self.h.phase_similarity(self.wc, cfg.ni, self.nw1, self.phase, cfg.nk, self.wk)

    def sigc(self):
        cfg = self.cfg
        ni, nb, ntq, nk = cfg.ni, cfg.nb, cfg.nq, cfg.nk
        for k in range(nk):
            m = int(self.nbt_k[k])
            self.h.zmel_dev(self._col(self.c_qmk, k, 0), self.c_q, self.zmel_sig.ptr + self.es * int(self.zoff[k]),
                            nm=m, nq=ntq, ndima=nb, nbatch=1, is_m_basis=True, zmelconjg=False)
            self.h.sigc_weights_dev(1, cfg.nf, m, ntq, self.d_omega, self.d_ekc.ptr + 8 * int(self.eoff[k]),
                                    cfg.esmr, cfg.ua, self.d_wk.ptr + 8 * k, self.d_x, self.d_wx, self.d_expa,
                                    self.wgt.ptr + 8 * int(self.goff[k]))
        self.zsec.zero()
        self.h.sigc_grouped_dev(self.nbt_k, self.zmel_sig, ni, ntq, cfg.nf, self.wk, self.wgt, self.zsec)

    def step(self):
        self.chi0()
        self.wbuild()
        self.sigc()

    def capture_step(self):
        self.h.graph_begin()
        try:
            self.step()
        finally:
            g = self.h.graph_end()
        return g

    def results(self):
        self.h.sync()
        ntq = self.cfg.nq
        chi, wc, zsec = self.chi.download(), self.wc.download(), self.zsec.download()
        sex, sexcore, vxc = synth.sigma_inputs(self.qp)
        return chi, wc, zsec, qp_energy(self.qp.e_q[:ntq], sex, sexcore, zsec, vxc)


def flops_uneven(qp: synth.QPoint, npb: int) -> float:
    """Algorithmic FLOPs of one uneven q-point (conventions of qsgw.flops_per_q)."""
    cfg = qp.cfg
    no_k, nu_k = synth.band_windows(qp)
    nbt_k = no_k + nu_k
    ni, nb, nw1, ntq, na = cfg.ni, cfg.nb, cfg.nf + 1, cfg.nq, cfg.natom
    f = 0.0
    f += float(np.sum(4.0 * nb * nb * ni * no_k / na ** 2 + 8.0 * ni * no_k * nu_k * nb / na))
    f += 8.0 * ni * ni * npb
    f += nw1 * 5 * 8.0 * ni ** 3
    f += float(np.sum(4.0 * nb * nb * ni * nbt_k / na ** 2 + 8.0 * ni * nbt_k * ntq * nb / na))
    f += float(np.sum(8.0 * ni * ni * nw1 * nbt_k * ntq + 8.0 * ntq * ntq * ni * nbt_k))
    return f
