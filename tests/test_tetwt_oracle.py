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


# ---------------------------------------------------------------- npm = 2 and core states
def _segments(r, jpm=None):
    """{(n1b, n2b): (ihw, whw segment)} of every k (1-based bands)."""
    pick = (lambda a: a) if jpm is None else (lambda a: a[..., jpm])
    out = {}
    nbnb = pick(r["nbnb"])
    for k in range(len(nbnb)):
        for i in range(nbnb[k]):
            j0, nh = pick(r["jhw"])[i, k] - 1, pick(r["nhw"])[i, k]
            out[(k, pick(r["n1b"])[i, k], pick(r["n2b"])[i, k])] = (pick(r["ihw"])[i, k], r["whw"][j0:j0 + nh])
    return out


def _assert_same_segments(a, b, rtol=0.0):
    assert a.keys() == b.keys()
    for key, (ih, w) in a.items():
        ih2, w2 = b[key]
        assert ih == ih2 and len(w) == len(w2), key
        if rtol == 0.0:
            np.testing.assert_array_equal(w, w2, err_msg=str(key))
        else:
            np.testing.assert_allclose(w, w2, rtol=rtol, atol=0.0, err_msg=str(key))


@pytest.mark.parametrize("ebmx", [1e10, 2.5])
def test_tetwt_npm2_first_half_is_npm1(orc, ebmx):
    """npm = 2: the jpm = 1 lists, pointers and weights are those of npm = 1
    (the whw pointers run over jpm first)."""
    m, ek1, ek2, ef, fr = _case(seed=4)
    r1 = orc.tetwt(m, ek1, ek2, ef, fr, ebmx=ebmx)
    r2 = orc.tetwt(m, ek1, ek2, ef, fr, ebmx=ebmx, npm=2)
    assert r2["nbnb"].shape == (ek1.shape[1], 2)
    np.testing.assert_array_equal(r2["nbnb"][:, 0], r1["nbnb"])
    n = r1["nbnbx"]
    for key in ("n1b", "n2b", "ihw", "nhw", "jhw"):
        np.testing.assert_array_equal(r2[key][:n, :, 0], r1[key][:n], err_msg=key)
    np.testing.assert_array_equal(r2["whw"][:r1["nhwtot"]], r1["whw"])
    assert r2["nhwtot"] > r1["nhwtot"]
    # the jpm = 2 pointers continue after the jpm = 1 weights
    assert r2["jhw"][0, 0, 1] == r1["nhwtot"] + 1


@pytest.mark.parametrize("ebmx", [1e10, 2.5])
def test_tetwt_npm2_second_half_is_swapped_npm1(orc, ebmx):
    """jpm = 2 pairs a state n at k with an occupied m at k+q; it is npm = 1
    with the k and k+q energies exchanged and each pair (n, m) read as (m, n):
    the same pairs, windows and histograms (ib-major order on each side)."""
    m, ek1, ek2, ef, fr = _case(seed=5)
    r2 = orc.tetwt(m, ek1, ek2, ef, fr, ebmx=ebmx, npm=2)
    rs = orc.tetwt(m, ek2, ek1, ef, fr, ebmx=ebmx)
    a = _segments(r2, 1)
    b = {(k, jb, ib): v for (k, ib, jb), v in _segments(rs).items()}
    assert len(a) > 0
    _assert_same_segments(a, b, rtol=1e-15)
    for k in range(ek1.shape[1]):          # ib-major: sorted by (n1b, n2b)
        n = r2["nbnb"][k, 1]
        pairs = list(zip(r2["n1b"][:n, k, 1], r2["n2b"][:n, k, 1]))
        assert pairs == sorted(pairs)


def _with_flat_bands(ek, ecore):
    nq = ek.shape[1]
    return np.vstack([np.repeat(np.asarray(ecore)[:, None], nq, axis=1), ek])


@pytest.mark.parametrize("npm", [1, 2])
def test_tetwt_core_is_a_flat_band(orc, npm):
    """Core states (appended after the valence bands, flat in k) give the
    weights of flat valence bands of the same energies placed below the
    valence bands; a degenerate core pair is averaged like a degenerate band."""
    m, ek1, ek2, ef, _ = _case(seed=6)
    ecore = np.array([-32.0, -32.0, -19.5]) / synth.RYDBERG_EV
    nband, nc = ek1.shape[0], len(ecore)
    f1, f2 = _with_flat_bands(ek1, ecore), _with_flat_bands(ek2, ecore)
    fr = synth.tetra_frhis(f1, f2)
    rc = orc.tetwt(m, ek1, ek2, ef, fr, npm=npm, ecore=ecore)
    rv = orc.tetwt(m, f1, f2, ef, fr, npm=npm)

    def band(b):  # core-form band (1-based) -> flat-band form
        return b - nband if b > nband else b + nc

    for j in range(npm):
        jj = None if npm == 1 else j
        sc = {(k, band(a), band(b)): v for (k, a, b), v in _segments(rc, jj).items()}
        sv = _segments(rv, jj)
        _assert_same_segments(sc, sv)
        core_side = [a for (_, a, b) in _segments(rc, jj)] if j == 0 else [b for (_, a, b) in _segments(rc, jj)]
        assert max(core_side) > nband      # core pairs are present
    np.testing.assert_array_equal(np.sort(rc["whw"]), np.sort(rv["whw"]))
    # the degenerate core pair carries equal histograms
    seg = _segments(rc, None if npm == 1 else 0)
    both = [(k, b) for (k, a, b) in seg if a == nband + 1]
    assert both
    for k, b in both:
        np.testing.assert_array_equal(seg[(k, nband + 1, b)][1], seg[(k, nband + 2, b)][1])


# ---------------------------------------------------------------- tetrahedron_matrix_linear
# With matrix_linear each bin's weight (times 4) is split over the four main
# corners as the barycentric coordinates of one point per integration
# tetrahedron: the vertex average of its cross-section polygon at the bin
# centre (mkwcg), clamped to the lowest / highest corner outside its range.

def _ml_tet(rng, no, nu, ef=0.0):
    kv = rng.normal(size=(4, 3))
    ea = rng.permutation(np.concatenate([rng.uniform(-1.0, -0.05, no), rng.uniform(0.05, 0.7, 4 - no)]))
    eb = rng.permutation(np.concatenate([rng.uniform(0.05, 1.0, nu), rng.uniform(-0.7, -0.05, 4 - nu)]))
    return kv, ea + ef, eb + ef


def _all_figures(seed, reps=3):
    rng = np.random.default_rng(seed)
    for t in range(25 * reps):
        yield (t,) + _ml_tet(rng, t % 5, (t // 5) % 5)


def test_lindtet6_ml_corner_sum_is_four_times_plain(orc):
    fr = orc.frhis(1.6)[1]
    for t, kv, ea, eb in _all_figures(11):
        x = 0.5 * (ea - eb)
        plain = orc.lindtet6(kv, ea, eb, x, 0.0, 0.0, fr)
        ml = orc.lindtet6_ml(kv, ea, eb, x, 0.0, 0.0, fr)
        assert ml.shape == (len(fr) - 1, 4)
        np.testing.assert_allclose(ml.sum(axis=1), 4.0 * plain, rtol=1e-12, atol=1e-15 * plain.sum(), err_msg=t)


def _cross_section_average(w, ec):
    """Barycentric coordinates of the vertex average of {w = ec} over a
    tetrahedron (w at its corners): brute force over the six edges."""
    lo, hi = int(np.argmin(w)), int(np.argmax(w))
    if ec <= w[lo]:
        return np.eye(4)[lo]
    if ec >= w[hi]:
        return np.eye(4)[hi]
    pts = []
    for i in range(4):
        for j in range(i + 1, 4):
            if (w[i] - ec) * (w[j] - ec) < 0.0:
                t = (ec - w[i]) / (w[j] - w[i])
                p = np.zeros(4)
                p[i], p[j] = 1.0 - t, t
                pts.append(p)
    assert len(pts) in (3, 4)
    return np.mean(pts, axis=0)


def test_lindtet6_ml_uncut_tetrahedron_is_cross_section_average(orc):
    """ea < E_F < eb at every corner: one integration tetrahedron, the main
    one, so corner c of bin i carries 4 W_i times the c-th coordinate of the
    cross-section's vertex average at the bin centre."""
    rng = np.random.default_rng(12)
    fr = np.linspace(0.0, 1.5, 61)
    for t in range(40):
        kv = rng.normal(size=(4, 3))
        ea = rng.uniform(-1.0, -0.05, 4)
        eb = rng.uniform(0.05, 1.0, 4)
        x = 0.5 * (ea - eb)
        w = -x
        plain = orc.lindtet6(kv, ea, eb, x, 0.0, 0.0, fr)
        ml = orc.lindtet6_ml(kv, ea, eb, x, 0.0, 0.0, fr)
        gap = np.diff(np.sort(w)).min()
        for i in np.nonzero(plain)[0]:
            lam = _cross_section_average(w, 0.5 * (fr[i] + fr[i + 1]))
            # the reference's corner offsets (up to 3e-8) move a coordinate by <= 3e-8 / (closest corner gap)
            np.testing.assert_allclose(ml[i], 4.0 * plain[i] * lam, rtol=0.0,
                                       atol=4.0 * plain[i] * 6e-8 / gap + 1e-14, err_msg=(t, i))


def test_lindtet6_ml_points_sit_on_the_bin_centre_inside_the_region(orc):
    """Every cut figure: the point that carries a bin lies in the tetrahedron
    (coordinates >= 0), in the occupied / unoccupied region (ea <= E_F <= eb
    there, both linear), and on the excitation plane w = bin centre, up to
    half a bin where a piece's range starts or ends inside the bin."""
    fr = np.linspace(0.0, 1.2, 2401)
    h = fr[1] - fr[0]
    for t, kv, ea, eb in _all_figures(13):
        x = 0.5 * (ea - eb)
        plain = orc.lindtet6(kv, ea, eb, x, 0.0, 0.0, fr)
        ml = orc.lindtet6_ml(kv, ea, eb, x, 0.0, 0.0, fr)
        tot = plain.sum()
        if tot == 0.0:
            assert not ml.any()
            continue
        m = 4.0 * plain
        assert ml.min() >= -1e-12 * tot, t
        ec = 0.5 * (fr[1:] + fr[:-1])
        assert np.all(ml @ (-x) - m * ec <= m * (0.5 * h + 1e-7) + 1e-14 * tot), t
        assert np.all(ml @ (-x) - m * ec >= -m * (0.5 * h + 1e-7) - 1e-14 * tot), t
        assert np.all(ml @ ea <= 1e-9 * m + 1e-14 * tot), t
        assert np.all(ml @ eb >= -1e-9 * m - 1e-14 * tot), t


def test_lindtet6_ml_affine_invariant_and_corner_covariant(orc):
    """The corner weights depend on k only through barycentric coordinates:
    unchanged under k -> A k + b; relabelling the corners relabels the
    columns (the integration tetrahedra follow the energies, not the labels)."""
    fr = orc.frhis(1.6)[1]
    rng = np.random.default_rng(14)
    for t, kv, ea, eb in _all_figures(15, reps=2):
        x = 0.5 * (ea - eb)
        ml = orc.lindtet6_ml(kv, ea, eb, x, 0.0, 0.0, fr)
        scale = np.abs(ml).max() + 1e-300
        a = rng.normal(size=(3, 3)) + 2.0 * np.eye(3)
        kv2 = kv @ a.T + rng.normal(size=3)
        det = abs(np.linalg.det(a))
        np.testing.assert_allclose(orc.lindtet6_ml(kv2, ea, eb, x, 0.0, 0.0, fr), det * ml,
                                   atol=1e-10 * det * scale, err_msg=t)
        p = rng.permutation(4)
        got = orc.lindtet6_ml(kv[p], ea[p], eb[p], x[p], 0.0, 0.0, fr)
        np.testing.assert_allclose(got, ml[:, p], atol=1e-10 * scale, err_msg=t)


def test_tetwt_ml_same_lists_and_bin_totals(orc):
    """tetrahedron_matrix_linear changes only whw: the same pair lists,
    windows and pointers, and per histogram bin the same total over all
    (k, pair) (each tetrahedron's 4 x W splits over its corners)."""
    m, ek1, ek2, ef, fr = _case(n=(3, 2, 2), seed=7)
    for npm in (1, 2):
        r0 = orc.tetwt(m, ek1, ek2, ef, fr, npm=npm)
        r1 = orc.tetwt(m, ek1, ek2, ef, fr, npm=npm, matrix_linear=True)
        for key in ("nbnb", "n1b", "n2b", "ihw", "nhw", "jhw"):
            np.testing.assert_array_equal(r1[key], r0[key], err_msg=key)
        nw = len(fr) - 1

        def bins(r):
            out = np.zeros(nw)
            seg = _segments(r, None) if npm == 1 else {**{(0,) + k: v for k, v in _segments(r, 0).items()},
                                                        **{(1,) + k: v for k, v in _segments(r, 1).items()}}
            for ih, w in seg.values():
                out[ih - 1:ih - 1 + len(w)] += w
            return out

        b0, b1 = bins(r0), bins(r1)
        np.testing.assert_allclose(b1, b0, rtol=1e-12, atol=1e-14 * b0.sum())
        assert not np.allclose(r1["whw"], r0["whw"], rtol=1e-3)   # the split is not the plain weight


def test_tetwt_ml_degenerate_bands_equal(orc):
    m = synth.tetra_mesh(2, 2, 2)
    ek1, ek2, ef = synth.tetra_bands(m, 4, 5, seed=3)
    ek1[2] = ek1[1]                 # a degenerate occupied pair at every k
    fr = synth.tetra_frhis(ek1, ek2)
    r = orc.tetwt(m, ek1, ek2, ef, fr, matrix_linear=True)
    seg = _segments(r)
    pairs = [(k, b) for (k, a, b) in seg if a == 2]
    assert pairs
    for k, b in pairs:
        np.testing.assert_array_equal(seg[(k, 2, b)][1], seg[(k, 3, b)][1])
