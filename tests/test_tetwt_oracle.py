"""Pins of the oracle's tetrahedron histogram weights (oracle/tetwt.c:
gettetwt + tetwt5x_dtet4 for npm=1).  No GPU.

The reference's tetwt5.f90 needs its GWinput/MPI modules and cannot be
compiled standalone here, so the restatement is pinned by what the method
computes, with independent numerics:

* lindtet6 = int_bin d omega int_tet d^3k f(ea)(1-f(eb)) delta(omega + x):
  Monte Carlo over uniform points of the tetrahedron (linear ea, eb, x),
  for every pair of occupied / unoccupied cut figures;
* the pair lists, bin windows and pointers of every k: brute force over the
  tetrahedra with numpy;
* the whole-mesh histogram of each (k, pair): pi/(4 vol_BZ) x the Monte
  Carlo histograms of the tetrahedra at k;
* degenerate bands: equal weights (the reference's symmetrisation).
"""
import numpy as np
import pytest

from ecalj_amd import synth


def _mc_hist(kv, ea, eb, x, ef, fr, n, rng):
    """Monte Carlo of lindtet6's integral over one tetrahedron."""
    lam = rng.dirichlet(np.ones(4), n)                       # uniform in the simplex
    vol = abs(np.linalg.det(np.array([kv[i] - kv[3] for i in range(3)]))) / 6.0
    a, b, xx = lam @ ea, lam @ eb, lam @ x
    w = np.where((a < ef) & (b > ef), vol / n, 0.0)
    h, _ = np.histogram(-xx, bins=fr, weights=w)
    return h, vol


def _figure(e, ef, occ):
    return int(np.sum(e < ef)) if occ else int(np.sum(e > ef))


def test_lindtet6_monte_carlo_every_figure(orc):
    rng = np.random.default_rng(5)
    fr = orc.frhis(1.6)[1]
    seen = set()
    for t in range(75):
        kv = rng.normal(size=(4, 3))
        no, nu = t % 5, (t // 5) % 5   # corners with ea < E_F, corners with eb > E_F
        ea = rng.permutation(np.concatenate([rng.uniform(-1.0, -0.05, no), rng.uniform(0.05, 0.7, 4 - no)]))
        eb = rng.permutation(np.concatenate([rng.uniform(0.05, 1.0, nu), rng.uniform(-0.7, -0.05, 4 - nu)]))
        x = 0.5 * (ea - eb)
        seen.add((_figure(ea, 0.0, True), _figure(eb, 0.0, False)))
        got = orc.lindtet6(kv, ea, eb, x, 0.0, 0.0, fr)
        ref, vol = _mc_hist(kv, ea, eb, x, 0.0, fr, 400000, rng)
        # cumulative distributions: MC error ~ vol / sqrt(N)
        assert np.abs(np.cumsum(got) - np.cumsum(ref)).max() < 6e-3 * vol, t
        assert abs(got.sum() - ref.sum()) < 6e-3 * vol
    # occupied corners 0..4 x unoccupied corners 0..4 (the cut figures) all exercised
    assert len(seen) == 25


def test_lindtet6_fully_inside_is_volume_distribution(orc):
    """ea < E_F and eb > E_F at all corners: the histogram is the distribution
    of -x over the tetrahedron, summing to its volume (Rath & Freeman)."""
    rng = np.random.default_rng(6)
    fr = orc.frhis(2.0)[1]
    kv = rng.normal(size=(4, 3))
    ea = rng.uniform(-1.0, -0.1, 4)
    eb = rng.uniform(0.2, 1.5, 4)
    got = orc.lindtet6(kv, ea, eb, 0.5 * (ea - eb), 0.0, 0.0, fr)
    vol = abs(np.linalg.det(np.array([kv[i] - kv[3] for i in range(3)]))) / 6.0
    assert abs(got.sum() - vol) < 1e-12 * vol * 100
    ref, _ = _mc_hist(kv, ea, eb, 0.5 * (ea - eb), 0.0, fr, 400000, rng)
    assert np.abs(np.cumsum(got) - np.cumsum(ref)).max() < 6e-3 * vol


def _case(n=(2, 2, 2), nocc=4, nun=5, seed=1):
    m = synth.tetra_mesh(*n)
    ek1, ek2, ef = synth.tetra_bands(m, nocc, nun, seed=seed)
    return m, ek1, ek2, ef, synth.tetra_frhis(ek1, ek2)


def test_tetwt_index_sets_bruteforce(orc):
    m, ek1, ek2, ef, fr = _case()
    r = orc.tetwt(m, ek1, ek2, ef, fr)
    nband, nqbz = ek1.shape
    flag = np.zeros((nband, nband, nqbz), bool)
    dmn = np.full((nband, nband, nqbz), 1e10)
    dmx = np.full((nband, nband, nqbz), -1e10)
    for t in range(m.idtetf.shape[1]):
        kk = m.ib1bz[m.idtetf[:, t] - 1] - 1
        nocc = max(int(np.sum(ek1[:, k] < ef)) for k in kk)
        nbmx = min(int(np.sum(ek2[:, k] < 1e10)) for k in kk)
        for ib in range(nocc):
            for jb in range(nbmx):
                eo, eu = ek1[ib, kk], ek2[jb, kk]
                if not (eo.min() <= ef and eu.max() >= ef) or (eu - eo).max() < 0:
                    continue
                d = 0.5 * (eu - eo)
                for k in kk:
                    flag[ib, jb, k] = True
                    dmn[ib, jb, k] = min(dmn[ib, jb, k], d.min())
                    dmx[ib, jb, k] = max(dmx[ib, jb, k], d.max())
    # degenerate groups (|de| <= 1e-4 between neighbours) share flags and windows
    for k in range(nqbz):
        g1 = np.concatenate([[0], np.cumsum(np.abs(np.diff(ek1[:, k])) > 1e-4)])
        g2 = np.concatenate([[0], np.cumsum(np.abs(np.diff(ek2[:, k])) > 1e-4)])
        for a in np.unique(g1):
            for b in np.unique(g2):
                ia, ib_ = np.where(g1 == a)[0], np.where(g2 == b)[0]
                blk = np.ix_(ia, ib_, [k])
                f, lo, hi = flag[blk].any(), dmn[blk].min(), dmx[blk].max()
                flag[blk], dmn[blk], dmx[blk] = f, lo, hi
    jtot = 1
    for k in range(nqbz):
        ib, jb = np.nonzero(flag[:, :, k])          # row-major: ib outer, jb inner (rsvwwk00_4)
        n = len(ib)
        assert r["nbnb"][k] == n
        np.testing.assert_array_equal(r["n1b"][:n, k], ib + 1)
        np.testing.assert_array_equal(r["n2b"][:n, k], jb + 1)
        for i in range(n):
            lo, hi = dmn[ib[i], jb[i], k], dmx[ib[i], jb[i], k]
            first = int(np.argmax(lo < fr[1:])) + 1          # first bin whose right edge exceeds demin
            last = int(np.argmax(hi < fr[1:])) + 1
            assert r["ihw"][i, k] == first and r["nhw"][i, k] == last - first + 1
            assert r["jhw"][i, k] == jtot
            jtot += last - first + 1
    assert r["nhwtot"] == jtot - 1


def test_tetwt_histograms_monte_carlo(orc):
    rng = np.random.default_rng(8)
    m, ek1, ek2, ef, fr = _case(nocc=3, nun=3, seed=2)
    r = orc.tetwt(m, ek1, ek2, ef, fr)
    nband, nqbz = ek1.shape
    voltot = abs(np.linalg.det(m.qlat))
    ref = {}
    for t in range(m.idtetf.shape[1]):
        w = m.idtetf[:, t] - 1
        kk = m.ib1bz[w] - 1
        kv = m.qbzw[:, w].T
        nocc = max(int(np.sum(ek1[:, k] < ef)) for k in kk)
        for ib in range(nocc):
            for jb in range(nband):
                eo, eu = ek1[ib, kk], ek2[jb, kk]
                if not (eo.min() <= ef and eu.max() >= ef) or (eu - eo).max() < 0:
                    continue
                h, _ = _mc_hist(kv, eo, eu, 0.5 * (eo - eu), ef, fr, 60000, rng)
                for k in kk:
                    ref[(k, ib, jb)] = ref.get((k, ib, jb), 0) + h * np.pi / (4 * voltot)
    tot_ref = sum(v.sum() for v in ref.values())
    assert abs(r["whw"].sum() - tot_ref) < 5e-3 * tot_ref
    for k in range(nqbz):
        for i in range(r["nbnb"][k]):
            ib, jb = r["n1b"][i, k] - 1, r["n2b"][i, k] - 1
            seg = np.zeros(len(fr) - 1)
            seg[r["ihw"][i, k] - 1:r["ihw"][i, k] - 1 + r["nhw"][i, k]] = \
                r["whw"][r["jhw"][i, k] - 1:r["jhw"][i, k] - 1 + r["nhw"][i, k]]
            h = ref.get((k, ib, jb))
            if h is None:  # a degenerate partner: symmetrised weights only
                continue
            if ib in (0, 1) and ek1[0, k] == ek1[1, k]:  # a doubled band: compare the pair average
                h = 0.5 * (ref.get((k, 0, jb), 0) + ref.get((k, 1, jb), 0))
            scale = max(np.abs(h).sum(), 1e-3)
            assert np.abs(np.cumsum(seg) - np.cumsum(h)).max() < 2e-2 * scale, (k, ib, jb)


def test_tetwt_degenerate_bands_equal(orc):
    m, ek1, ek2, ef, fr = _case(nocc=5, nun=4, seed=3)
    assert np.abs(ek1[0] - ek1[1]).max() == 0.0   # the doubled band
    r = orc.tetwt(m, ek1, ek2, ef, fr)
    for k in range(ek1.shape[1]):
        n = r["nbnb"][k]
        seg = {}
        for i in range(n):
            j0, nh = r["jhw"][i, k] - 1, r["nhw"][i, k]
            seg[(r["n1b"][i, k], r["n2b"][i, k])] = (r["ihw"][i, k], r["whw"][j0:j0 + nh])
        for (a, b), (ih, w) in seg.items():
            if a == 1:
                ih2, w2 = seg[(2, b)]
                assert ih2 == ih
                np.testing.assert_array_equal(w, w2)


def test_hisrange_and_chkdgn(orc):
    fr = np.array([0.0, 0.1, 0.25, 0.5, 1.0])
    assert orc.lib().orc_hisrange is not None
    a, b = orc.chkdgn(np.array([0.0, 0.00005, 0.3, 0.3, 0.7]))
    np.testing.assert_array_equal(a, [1, 3, 5])
    np.testing.assert_array_equal(b, [2, 4, 5])
