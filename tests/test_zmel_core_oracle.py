"""Pins of the oracle's zmel with core states and the rotated-atom map
(orc_zmel_init, restating S/m_zmel.f90:199-214, 323-391) by what they mean,
through the valence-only zmel (orc_zmel, pinned by numpy einsum elsewhere):

* no core states and the identity map: orc_zmel_init = orc_zmel;
* a core state is a state whose muffin-tin coefficient vector is the unit
  vector of its core function: with every atom's functions extended to
  core + valence (the full ppb block) and the eigenvectors padded with zeros
  on the core functions, the valence formula gives the core rows too;
* iatomp: atom ia takes the product basis, ppb block and rows of atom
  iatomp(ia), which equals the identity map on ppb blocks permuted by
  iatomp, with the product-basis rows moved back to the places of iatomp."""
import numpy as np
import pytest


def _case(rng, natom=3, nc=(2, 0, 1), nv=(5, 4, 6), md=(7, 6, 8), nlnmx=None, mdimx=None):
    nc, nv, md = list(nc), list(nv), list(md)
    nlnmx = nlnmx or max(c + v for c, v in zip(nc, nv)) + 1
    mdimx = mdimx or max(md) + 1
    ppb = np.zeros((nlnmx, nlnmx, mdimx, natom), order="F")
    for a in range(natom):
        n = nc[a] + nv[a]
        ppb[:n, :n, :md[a], a] = rng.standard_normal((n, n, md[a]))
    ndima = sum(nv) + 2
    return nc, nv, md, ppb, ndima


def _crand(rng, *s):
    return np.asfortranarray(rng.standard_normal(s) + 1j * rng.standard_normal(s))


def test_valence_only_equals_orc_zmel(orc):
    rng = np.random.default_rng(1)
    nc, nv, md, ppb, ndima = _case(rng, nc=(0, 0, 0))
    cm, cq = _crand(rng, ndima, 4), _crand(rng, ndima, 3)
    ph = np.exp(1j * rng.uniform(0, 6, 3))
    blocks = np.concatenate([ppb[:nv[a], :nv[a], :md[a], a].ravel(order="F") for a in range(3)])
    ref = orc.zmel(nv, md, ph, blocks, cm, cq, conjg=True)
    got = orc.zmel_init(nc, nv, md, ppb, 1, 4, 0, cm, cq, phase=ph, conjg=True)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("ns1,ns2", [(1, 6), (2, 3), (4, 7), (1, 3)])
def test_core_rows_are_the_valence_formula_on_unit_vectors(orc, ns1, ns2):
    """nctot = 3 core states (2 on atom 1, 1 on atom 3), then valence bands."""
    rng = np.random.default_rng(ns1 * 10 + ns2)
    nc, nv, md, ppb, ndima = _case(rng)
    nctot = sum(nc)
    nmv = max(0, ns2 - max(ns1, nctot + 1) + 1)
    cm, cq = _crand(rng, ndima, nmv), _crand(rng, ndima, 4)
    ph = np.exp(1j * rng.uniform(0, 6, 3))
    got = orc.zmel_init(nc, nv, md, ppb, ns1, ns2, nctot, cm, cq, phase=ph)
    # extended per-atom function spaces (core + valence), rows ordered atom by atom
    ne = [c + v for c, v in zip(nc, nv)]
    off_e = np.concatenate([[0], np.cumsum(ne)])
    off_v = np.concatenate([[0], np.cumsum(nv)])
    off_c = np.concatenate([[0], np.cumsum(nc)])
    blocks = np.concatenate([ppb[:ne[a], :ne[a], :md[a], a].ravel(order="F") for a in range(3)])
    nre = int(off_e[-1])
    cq_e = np.zeros((nre, cq.shape[1]), complex)
    for a in range(3):
        cq_e[off_e[a] + nc[a]:off_e[a + 1]] = cq[off_v[a]:off_v[a + 1]]
    cols = []
    for it in range(ns1, ns2 + 1):
        col = np.zeros(nre, complex)
        if it <= nctot:  # core state it: unit vector of its core function on its atom
            a = int(np.searchsorted(off_c, it, side="left")) - 1
            col[off_e[a] + (it - 1 - off_c[a])] = 1.0
        else:            # valence band it - nctot: column of cphim
            j = it - max(ns1, nctot + 1)
            for a in range(3):
                col[off_e[a] + nc[a]:off_e[a + 1]] = cm[off_v[a]:off_v[a + 1], j]
        cols.append(col)
    cm_e = np.asfortranarray(np.stack(cols, axis=1))
    ref = orc.zmel(ne, md, ph, blocks, cm_e, np.asfortranarray(cq_e))
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()


def test_iatomp_is_a_permutation_of_blocks_and_rows(orc):
    rng = np.random.default_rng(7)
    # equivalent atoms (same sizes) so that any permutation is a valid map
    nc, nv, md, ppb, ndima = _case(rng, natom=3, nc=(1, 1, 1), nv=(4, 4, 4), md=(5, 5, 5))
    nctot = 3
    cm, cq = _crand(rng, ndima, 3), _crand(rng, ndima, 2)
    ph = np.exp(1j * rng.uniform(0, 6, 3))
    perm = np.array([2, 3, 1])          # iatomp, 1-based
    got = orc.zmel_init(nc, nv, md, ppb, 1, 6, nctot, cm, cq, phase=ph, iatomp=perm)
    ppb_p = np.asfortranarray(ppb[:, :, :, perm - 1])
    ident = orc.zmel_init(nc, nv, md, ppb_p, 1, 6, nctot, cm, cq, phase=ph)
    ref = np.zeros_like(ident)
    for a in range(3):    # rows of atom a in the identity run sit at atom perm(a)'s rows
        ap = perm[a] - 1
        ref[5 * ap:5 * ap + 5] = ident[5 * a:5 * a + 5]
    assert np.abs(got - ref).max() < 1e-13 * np.abs(ref).max()
