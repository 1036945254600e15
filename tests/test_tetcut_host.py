"""The per-tetrahedron integration of the HIP tetrahedron kernel
(ecalj_amd/csrc/tetcut.hpp, compiled here for the host with g++) against the
oracle's restatement of lindtet6 (pinned by Monte Carlo in
test_tetwt_oracle.py): every cut figure, equal corner energies, corners on
E_F, and every (tetrahedron, pair) of the meshes of the GPU test; the same
with tetrahedron_matrix_linear's per-corner shares.

Bars.  Against the same pieces evaluated in 50-digit arithmetic (mpmath),
every bin within 1e-13 of the tetrahedron's volume.  Against the oracle,
every bin within 5e-8 of the volume and every total within 1e-12 of it: the
oracle forms the middle branch of the distribution as (a - b)(a^2 + ab + b^2)
with a, b cube roots, as the reference does, and where the 1e-8 corner
offsets separate two equal energies that form cancels terms of size
~(spread / 1e-8); measured, it is off by up to 1.5e-8 of the volume, where the
kernel's form (Bloechl et al.) stays within 5e-16."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from ecalj_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def _cutlib(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("tetcut") / "libtetcut.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", "-shared", "-fPIC",
                    "-I", os.path.join(ROOT, "ecalj_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "tetcut_host.cpp"), "-o", so], check=True)
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    for f in (lib.tetcut_pair_histogram, lib.tetcut_pair_histogram_ml):
        f.argtypes = [dp, dp, dp, C.c_double, dp, C.c_int, dp]
        f.restype = C.c_int

    def run(f, ncol):
        def call(kvec, ea, eb, ef, fr):
            k = np.ascontiguousarray(kvec, np.float64)
            ea, eb, fr = (np.ascontiguousarray(a, np.float64) for a in (ea, eb, fr))
            out = np.zeros((len(fr) - 1, ncol))
            rc = f(k.ctypes.data_as(dp), ea.ctypes.data_as(dp), eb.ctypes.data_as(dp), ef, fr.ctypes.data_as(dp),
                   len(fr) - 1, out.ctypes.data_as(dp))
            if rc:
                raise ValueError(rc)
            return out[:, 0] if ncol == 1 else out
        return call
    return run(lib.tetcut_pair_histogram, 1), run(lib.tetcut_pair_histogram_ml, 4)


@pytest.fixture(scope="module")
def cut(_cutlib):
    return _cutlib[0]


@pytest.fixture(scope="module")
def cut_ml(_cutlib):
    """tetrahedron_matrix_linear: nwhis x 4, column c carried by corner c."""
    return _cutlib[1]


def _vol(kvec):
    k = np.asarray(kvec, np.float64)
    return abs(np.linalg.det(k[1:] - k[0])) / 6.0


def _same(a, b, kvec):
    assert np.abs(a - b).max() <= 5e-8 * _vol(kvec)
    assert abs(a.sum() - b.sum()) <= 1e-12 * _vol(kvec)


def _oracle(orc, kvec, ea, eb, ef, fr):
    return orc.lindtet6(kvec, ea, eb, 0.5 * (np.asarray(ea) - np.asarray(eb)), ef, ef, fr)


# the clip tables of tetcut.hpp: vertex (a, b) is corner a, or the zero of the
# clip value on edge a-b; corners labelled by rank, the n inside last
_CLIP = {0: [], 1: [[(3, 3), (3, 1), (3, 0), (3, 2)]],
         2: [[(3, 3), (2, 2), (3, 1), (3, 0)], [(2, 2), (3, 0), (2, 0), (3, 1)], [(2, 2), (3, 1), (2, 0), (2, 1)]],
         3: [[(2, 2), (3, 3), (2, 0), (1, 1)], [(3, 3), (3, 0), (1, 0), (2, 0)], [(3, 3), (1, 1), (1, 0), (2, 0)]],
         4: [[(0, 0), (1, 1), (2, 2), (3, 3)]]}


def _exact(kvec, ea, eb, ef, fr):
    """The same pieces and offsets in 50-digit arithmetic (mpmath); the
    distribution of w over a piece in the two-term form
    (t-s0)^3/(d10 d20 d30) - (t-s1)^3/(d10 d21 d31) on its middle branch."""
    import mpmath as mp

    mp.mp.dps = 50
    f = [mp.mpf(float(x)) for x in fr]
    ef = mp.mpf(float(ef))

    def rank(key, later_first):
        lab = [0] * 4
        for c in range(4):
            lab[sum(1 for d in range(4) if d != c and (key[d] < key[c] or (key[d] == key[c] and
                                                                            (d > c if later_first else d < c))))] = c
        return lab

    def vert(lab, g, v):
        a, b = v
        if a == b:
            return lab[a]
        t = g[a] / (g[a] - g[b])
        return tuple(lab[a][i] + t * (lab[b][i] - lab[a][i]) for i in range(5))

    out = [mp.mpf(0)] * (len(f) - 1)
    ea = [mp.mpf(float(x)) for x in ea]
    eb = [mp.mpf(float(x)) for x in eb]
    order = rank([-x for x in ea], True)
    lab = [tuple(mp.mpf(float(x)) for x in kvec[c]) + ((eb[c] - ea[c]) / 2, eb[c] - ef) for c in order]
    g = [ef - ea[c] for c in order]
    for pc in _CLIP[sum(1 for x in g if x > 0)]:
        v = [vert(lab, g, x) for x in pc]
        lab2 = [v[i] for i in rank([x[4] for x in v], False)]
        g2 = [x[4] for x in lab2]
        for pc2 in _CLIP[sum(1 for x in g2 if x > 0)]:
            q = [vert(lab2, g2, x) for x in pc2]
            vol = abs(mp.det(mp.matrix([[q[i][j] - q[0][j] for j in range(3)] for i in range(1, 4)]))) / 6
            s = sorted(x[3] for x in q)
            s = [s[0] - mp.mpf("3e-8"), s[1] - mp.mpf("2e-8"), s[2] - mp.mpf("1e-8"), s[3]]

            def cum(t):
                if t < s[0]:
                    return mp.mpf(0)
                if t >= s[3]:
                    return mp.mpf(1)
                lo = (t - s[0]) ** 3 / ((s[1] - s[0]) * (s[2] - s[0]) * (s[3] - s[0]))
                if t < s[1]:
                    return lo
                if t < s[2]:
                    return lo - (t - s[1]) ** 3 / ((s[1] - s[0]) * (s[2] - s[1]) * (s[3] - s[1]))
                return 1 - (s[3] - t) ** 3 / ((s[3] - s[0]) * (s[3] - s[1]) * (s[3] - s[2]))
            below = mp.mpf(0)
            for i in range(1, len(f)):
                now = cum(f[i])
                out[i - 1] += vol * (now - below)
                below = now
    return np.array([float(x) for x in out])


FR = np.concatenate([[0.0], np.geomspace(1e-3, 4.0, 160)])


def _figures(rng, n):
    """Random figures on tetrahedra of a mesh-like shape (no near-flat ones,
    whose volume is below the rounding of its corner coordinates)."""
    for _ in range(n):
        kvec = rng.uniform(-1, 1, (4, 3))
        if _vol(kvec) < 0.02:
            continue
        ea = rng.uniform(-1.0, 0.6, 4)
        eb = rng.uniform(-0.6, 1.5, 4)
        if (eb - ea).max() >= 0:
            yield kvec, ea, eb


def test_every_figure(cut, orc):
    """1..4 occupied corners x 1..4 unoccupied corners, random tetrahedra."""
    seen = set()
    for kvec, ea, eb in _figures(np.random.default_rng(3), 4000):
        seen.add((int((ea < 0).sum()), int((eb > 0).sum())))
        _same(cut(kvec, ea, eb, 0.0, FR), _oracle(orc, kvec, ea, eb, 0.0, FR), kvec)
    assert {(a, b) for a in range(1, 5) for b in range(1, 5)} <= seen


def test_against_50_digits(cut):
    """The kernel's arithmetic against the same integration in 50 digits,
    on random figures, on equal corner energies of one band, and on a flat
    excitation (all four values of w pulled apart only by the offsets).
    Equal energies of both bands are left to the oracle comparison: there a
    cut point can tie with a corner exactly, and the ranking of such a tie
    follows the rounding of the cut point."""
    rng = np.random.default_rng(5)
    lv = np.array([-0.4, -0.1, 0.0, 0.2])
    cases = list(_figures(rng, 60))
    while len(cases) < 150:
        kvec = rng.uniform(-1, 1, (4, 3))
        if _vol(kvec) < 0.02:
            continue
        ea, eb = rng.uniform(-0.6, 0.3, 4), rng.uniform(-0.3, 0.8, 4)
        kind = len(cases) % 3
        if kind == 0:
            ea = rng.choice(lv, 4)
        elif kind == 1:
            eb = rng.choice(lv + 0.3, 4)
        else:
            eb = ea + 0.25
        cases.append((kvec, ea, eb))
    for kvec, ea, eb in cases:
        assert np.abs(cut(kvec, ea, eb, 0.0, FR) - _exact(kvec, ea, eb, 0.0, FR)).max() <= 1e-13 * _vol(kvec)


def test_equal_energies(cut, orc):
    """Ties among corner energies (ranked as the reference ranks them),
    corners exactly on E_F, and a flat excitation."""
    rng = np.random.default_rng(4)
    lv = np.array([-0.4, -0.1, 0.0, 0.2])
    for _ in range(3000):
        kvec = rng.uniform(-1, 1, (4, 3))
        ea = rng.choice(lv, 4)
        eb = rng.choice(lv + 0.3, 4)
        if rng.random() < 0.3:
            eb = ea + 0.25
        _same(cut(kvec, ea, eb, 0.0, FR), _oracle(orc, kvec, ea, eb, 0.0, FR), kvec)


@pytest.mark.parametrize("n,nocc,nun,seed,degenerate", [((2, 2, 2), 4, 5, 1, True), ((3, 3, 3), 6, 20, 2, True),
                                                        ((3, 2, 4), 5, 12, 4, True)])
def test_mesh_pairs(cut, orc, n, nocc, nun, seed, degenerate):
    """Every (tetrahedron, band pair) of a symmetric mesh with smooth bands."""
    m = synth.tetra_mesh(*n)
    ek1, ek2, ef = synth.tetra_bands(m, nocc, nun, seed=seed, degenerate=degenerate)
    fr = synth.tetra_frhis(ek1, ek2)
    nb = ek1.shape[0]
    count = 0
    for t in range(m.idtetf.shape[1]):
        w = m.idtetf[:, t] - 1
        kk = m.ib1bz[w] - 1
        kvec = m.qbzw[:, w].T
        for ib in range(nb):
            ea = ek1[ib, kk]
            if ea.min() > ef:
                continue
            for jb in range(nb):
                eb = ek2[jb, kk]
                if eb.max() < ef or (eb - ea).max() < 0:
                    continue
                _same(cut(kvec, ea, eb, ef, fr), _oracle(orc, kvec, ea, eb, ef, fr), kvec)
                count += 1
    assert count > 100


def _same_ml(a, b, kvec):
    """matrix_linear: the bins carry 4 W, so the bars are 4 x those of the
    plain histogram; a corner share moves with its bin's weight."""
    assert np.abs(a - b).max() <= 2e-7 * _vol(kvec)
    assert abs(a.sum() - b.sum()) <= 4e-12 * _vol(kvec)


def test_matrix_linear_every_figure_and_ties(cut, cut_ml, orc):
    """Per-corner shares against the oracle (which finds each piece corner's
    coordinates by inverting the tetrahedron's k-vectors, as the reference
    does): random figures, equal corner energies, corners on E_F; and the
    corner sum is 4 x the plain histogram of the same code."""
    rng = np.random.default_rng(7)
    lv = np.array([-0.4, -0.1, 0.0, 0.2])
    cases = list(_figures(np.random.default_rng(8), 1500))
    for _ in range(800):
        kvec = rng.uniform(-1, 1, (4, 3))
        ea, eb = rng.choice(lv, 4), rng.choice(lv + 0.3, 4)
        if rng.random() < 0.3:
            eb = ea + 0.25
        cases.append((kvec, ea, eb))
    for kvec, ea, eb in cases:
        got = cut_ml(kvec, ea, eb, 0.0, FR)
        ref = orc.lindtet6_ml(kvec, ea, eb, 0.5 * (np.asarray(ea) - np.asarray(eb)), 0.0, 0.0, FR)
        _same_ml(got, ref, kvec)
        np.testing.assert_allclose(got.sum(axis=1), 4.0 * cut(kvec, ea, eb, 0.0, FR), rtol=1e-13,
                                   atol=1e-15 * _vol(kvec))


def test_matrix_linear_mesh_pairs(cut_ml, orc):
    m = synth.tetra_mesh(3, 2, 4)
    ek1, ek2, ef = synth.tetra_bands(m, 5, 12, seed=4, degenerate=True)
    fr = synth.tetra_frhis(ek1, ek2)
    nb = ek1.shape[0]
    count = 0
    for t in range(m.idtetf.shape[1]):
        w = m.idtetf[:, t] - 1
        kk = m.ib1bz[w] - 1
        kvec = m.qbzw[:, w].T
        for ib in range(nb):
            ea = ek1[ib, kk]
            if ea.min() > ef:
                continue
            for jb in range(nb):
                eb = ek2[jb, kk]
                if eb.max() < ef or (eb - ea).max() < 0:
                    continue
                ref = orc.lindtet6_ml(kvec, ea, eb, 0.5 * (ea - eb), ef, ef, fr)
                _same_ml(cut_ml(kvec, ea, eb, ef, fr), ref, kvec)
                count += 1
    assert count > 100
