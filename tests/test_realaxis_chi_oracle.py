"""Pins of the oracle's real-axis chi0 branch (orc_rmatt, orc_dpsion_real) and
of its pivoted inverse, which the real-axis W build needs.  No GPU.

* rmatt against hilbertmat compiled from the reference source (oracle/_ref);
* rmatt against principal-value quadrature of the piecewise-linear
  interpolant of the bins (scipy QAWC, an independent method);
* dpsion_real against the Kramers-Kronig pair of a two-pole response whose
  real part is known in closed form: binned Im chi in, Re chi out;
* the inverse against numpy.linalg.inv on matrices that need row exchanges.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate


def _edges(nwhis, fr):
    L = np.concatenate([-fr[nwhis:0:-1], [0.0], fr[0:nwhis]])
    R = np.concatenate([-fr[nwhis - 1::-1], [0.0], fr[1:nwhis + 1]])
    return L, 0.5 * (L + R), R


def _freqr(fr, nw):
    # getfreq (S/m_freq.f90:127-129): freq_r(0) = 0, freq_r(i) = bin centres
    return np.concatenate([[0.0], 0.5 * (fr[:nw] + fr[1:nw + 1])])


def test_rmatt_matches_reference_build(orc):
    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    fr = orc.frhis(0.6)[1]
    nwhis = len(fr) - 1
    nw = nwhis - 40
    freqr = _freqr(fr, nw)
    ours = orc.rmatt(freqr, fr)
    L, Cc, R = _edges(nwhis, fr)
    dp = C.POINTER(C.c_double)
    nn = C.c_int(nwhis)
    w_min = fr[1] - fr[0]
    for it in list(range(0, 12)) + list(range(nw - 5, nw + 1)) + [nw // 2]:
        out = np.zeros(2 * nwhis + 1, np.complex128)
        z = np.array([freqr[it]], np.complex128)
        ref.hilbertmat_(z.ctypes.data_as(dp), C.byref(nn), L.ctypes.data_as(dp), Cc.ctypes.data_as(dp),
                        R.ctypes.data_as(dp), out.ctypes.data_as(dp))
        # dpsion_init (S/dpsion5.f90:98-99): rmat(it,:) = Re(rrr)/pi; rmatt = rmat(iw) - rmat(-iw)
        want = (out[nwhis + 1:].real - out[nwhis - 1::-1].real) / math.pi
        tol = 1e-13 * (1.0 + freqr[it] / w_min)
        np.testing.assert_allclose(ours[it], want, rtol=0, atol=tol * np.abs(want).max())


def _interp_nodes(fr, fpos):
    """Nodes of the piecewise-linear interpolant hilbertmat integrates: odd
    extension of the bin values, zero at 0 and at the outer edges."""
    nwhis = len(fr) - 1
    L, Cc, R = _edges(nwhis, fr)
    f = np.concatenate([-fpos[::-1], [0.0], fpos])
    xs = np.concatenate([[L[0]], Cc[:nwhis], [0.0], Cc[nwhis + 1:], [R[-1]]])
    ys = np.concatenate([[0.0], f[:nwhis], [0.0], f[nwhis + 1:], [0.0]])
    return xs, ys


def _pv(xs, ys, w):
    """P int g(x)/(x-w) dx over [xs[0], xs[-1]], g the linear interpolant."""
    g = lambda x: np.interp(x, xs, ys)
    k = int(np.searchsorted(xs, w))  # xs[k-1] < w < xs[k]
    a, b = xs[k - 1], xs[k]
    total = integrate.quad(g, a, b, weight="cauchy", wvar=w, epsabs=1e-13, limit=200)[0]
    for lo, hi in ((xs[0], a), (b, xs[-1])):
        pts = [p for p in xs if lo < p < hi]
        total += integrate.quad(lambda x: g(x) / (x - w), lo, hi, points=pts[:1000] if len(pts) else None,
                                limit=4000, epsabs=1e-12)[0]
    return total


def test_rmatt_is_principal_value_of_linear_interp(orc):
    fr = orc.frhis(0.25)[1]
    nwhis = len(fr) - 1
    rng = np.random.default_rng(11)
    fpos = rng.standard_normal(nwhis)
    xs, ys = _interp_nodes(fr, fpos)
    # off-node points: inside bins, between the bin centres
    ws = [0.5 * (fr[5] + 0.5 * (fr[5] + fr[6])), 0.0371, 0.1234, 0.2011]
    rm = orc.rmatt(np.array(ws), fr)
    for i, w in enumerate(ws):
        lhs = rm[i] @ fpos
        rhs = _pv(xs, ys, w) / math.pi
        assert abs(lhs - rhs) < 1e-7 * max(1.0, abs(rhs)), (w, lhs, rhs)


def _two_pole(A, Om, gam, x):
    """chi(x) = A [1/(x - Om + i g) - 1/(x + Om + i g)]: retarded, chi(-x) = conj(chi(x))."""
    return A * (1.0 / (x - Om + 1j * gam) - 1.0 / (x + Om + 1j * gam))


def _kk_case(orc, A, Om, gam, ratio, npr_scale=(1.0,)):
    fr = orc.frhis(2.5, ratio)[1]
    nwhis = len(fr) - 1
    nw = int(np.searchsorted(fr, 1.2))
    freqr = _freqr(fr, nw)
    L, R = fr[:-1], fr[1:]
    # x0gemm output for this response: rcxq(bin) = -int_bin Im chi (positive sums of weights)
    at = lambda x, c: np.arctan((x - c) / gam)
    bins = A * ((at(R, Om) - at(L, Om)) - (at(R, -Om) - at(L, -Om)))
    sc = np.asarray(npr_scale)
    rcxq = np.zeros((len(sc), 1, nwhis + 1), np.complex128, order="F")
    rcxq[:, 0, 1:] = sc[:, None] * bins[None, :]
    rm = orc.rmatt(freqr, fr)
    zxq = orc.dpsion_real(rcxq, fr, rm)
    want = _two_pole(A, Om, gam, freqr)
    want[0] = want[0].real  # the zero-frequency slot carries no accumulated Im part
    return fr, freqr, bins, rcxq, rm, zxq, want


@pytest.mark.parametrize("Om,gam", [(0.2, 0.05), (0.5, 0.08)])
def test_dpsion_real_kramers_kronig(orc, Om, gam):
    """Binned Im chi in, chi(omega) out.  The reference's scheme treats bin
    averages as values at the bin centres of a geometric mesh, so Re chi
    carries an O(ratio-1) error and Im chi an O(width^2) one: check both
    bounds at the default mesh (1.03) and first-order convergence of Re chi
    as the mesh is refined.  A one-bin shift, a sign or a lost 1/pi breaks
    the bounds by orders of magnitude."""
    A = 0.7
    errs = []
    for ratio in (1.03, 1.015, 1.0075):
        fr, freqr, bins, rcxq, rm, zxq, want = _kk_case(orc, A, Om, gam, ratio, (1.0, -0.4 + 0.3j))
        got = zxq[0, 0]
        peak = np.abs(want).max()
        errs.append(np.abs(got.real - want.real).max() / peak)
        assert np.abs(got.imag - want.imag).max() < 5e-3 * peak * ((ratio - 1) / 0.03) ** 2
        assert got[0].imag == 0.0
        # linear in rcxq
        np.testing.assert_allclose(zxq[1, 0], (-0.4 + 0.3j) * got, rtol=1e-13, atol=1e-13 * peak)
    assert errs[0] < 0.05
    assert 1.7 < errs[0] / errs[1] < 2.3 and 1.7 < errs[1] / errs[2] < 2.3, errs
    # Re chi(0) converges to the closed form -2 A Om / (Om^2 + gam^2)
    assert abs(got[0].real - A * (-2 * Om) / (Om ** 2 + gam ** 2)) < 0.05 * peak
    # rcxq was scaled in place to the bin averages of Im chi (S/dpsion5.f90:193-200)
    L, R = fr[:-1], fr[1:]
    np.testing.assert_allclose(rcxq[0, 0, 1:], -bins / (R - L), rtol=1e-14)
    # scaled=True reuses the scaled array: the same zxq
    z2 = orc.dpsion_real(rcxq, fr, rm, scaled=True)
    np.testing.assert_allclose(z2, zxq, rtol=1e-14, atol=1e-14)


def test_dpsion_real_ecut_filter(orc):
    fr = orc.frhis(0.8)[1]
    nwhis = len(fr) - 1
    nw = nwhis // 2
    freqr = _freqr(fr, nw)
    rng = np.random.default_rng(2)
    r0 = np.asfortranarray(rng.standard_normal((2, 2, nwhis + 1)) + 1j * rng.standard_normal((2, 2, nwhis + 1)))
    rm = orc.rmatt(freqr, fr)
    ecut = 0.3
    z_cut = orc.dpsion_real(r0.copy(order="F"), fr, rm, ecut=ecut)
    hc = 0.5 * (fr[:-1] + fr[1:])
    r1 = r0.copy(order="F")
    r1[:, :, 1:] *= np.exp(-(hc / ecut) ** 2)
    z_ref = orc.dpsion_real(r1, fr, rm)
    # the slot-0 term is not filtered (the filter loop runs over bins 1..nwhis)
    z_ref[:, :, 0] += 1j * (r0[:, :, 0] - r1[:, :, 0])
    np.testing.assert_allclose(z_cut, z_ref, rtol=1e-13, atol=1e-13 * np.abs(z_ref).max())


@pytest.mark.parametrize("n", [7, 40, 101])
def test_zinv_needs_row_exchanges(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a[0, 0] = 0.0          # no-pivot elimination would divide by zero here
    a[1, :] = a[0, :] * 1e-12 + a[1, :] * 1e-9
    a = np.asfortranarray(a)
    inv = orc.zinv(a.copy(order="F"))
    ref = np.linalg.inv(a)
    assert np.abs(inv @ a - np.eye(n)).max() < 1e-6
    assert np.abs(inv - ref).max() / np.abs(ref).max() < 1e-6
