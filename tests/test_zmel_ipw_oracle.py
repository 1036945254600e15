"""Pins of the oracle's IPW block of zmel (orc_zmel_ipw, S/m_zmel.f90:418-538)
against a numpy triple sum over the two plane-wave bases, which takes no
shortcut through the combined-G tables:
  zmel(c, n, m) = phase(c) sum_{G1, G2} theta(G1 - Gc - G2 - nadd) geigq(G1, m) dgeigqk(G2, n),
theta = ggg where the table has the vector, 0 elsewhere.  Both contraction
orders of the reference (G1 first / G2 first) must give it.  No GPU."""
import numpy as np
import pytest

from ecalj_amd import synth


def _brute(c):
    lookup = {}
    lo = c.ggg_lo
    for idx in np.argwhere(c.igggi > 0):
        lookup[tuple(idx + lo)] = c.ggg[c.igggi[tuple(idx)] - 1]
    ngc = c.ngveccR.shape[1]
    g1, g2 = c.ngvecpB1.T, c.ngvecpB2.T
    th = np.zeros((ngc, len(g2), len(g1)), complex)
    for a in range(ngc):
        for j, v2 in enumerate(g2):
            for i, v1 in enumerate(g1):
                th[a, j, i] = lookup.get(tuple(v1 - c.ngveccR[:, a] - v2 - c.nadd), 0.0)
    geq = c.geigq[:len(g1)]
    dgk = c.dgeigqk[:len(g2)]
    return c.phase[:, None, None] * np.einsum("aji,im,jn->anm", th, geq, dgk)


@pytest.mark.parametrize("seed,ngc,ngp1,ngp2,nm,ntp0,nadd", [(1, 20, 15, 18, 7, 4, (0, 1, -1)),
                                                            (2, 27, 19, 14, 3, 6, (0, 0, 0)),
                                                            (3, 9, 7, 11, 5, 5, (1, -1, 2))])
def test_zmel_ipw_bruteforce_both_branches(orc, seed, ngc, ngp1, ngp2, nm, ntp0, nadd):
    c = synth.ipw_case(seed, ngc, ngp1, ngp2, nm, ntp0, nadd=nadd)
    ref = _brute(c)
    for b in (0, 1, 2):
        got = orc.zmel_ipw(c, branch=b)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.abs(ref).max() > 0.1


def test_zmel_ipw_linear_in_ggg(orc):
    """Scaling the step-function table scales the block; an empty table gives zero."""
    c = synth.ipw_case(4, 12, 10, 9, 4, 3)
    a = orc.zmel_ipw(c)
    c2 = synth.IPWCase(**{**c.__dict__, "ggg": 2.5j * c.ggg})
    np.testing.assert_allclose(orc.zmel_ipw(c2), 2.5j * a, rtol=1e-13, atol=1e-13)
    c3 = synth.IPWCase(**{**c.__dict__, "igggi": np.zeros_like(c.igggi)})
    assert np.abs(orc.zmel_ipw(c3)).max() == 0.0
