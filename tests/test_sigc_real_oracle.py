"""Pins of the real-axis Sigma_c oracle (orc_sigc_real, after
S/m_sxcf_gemm.f90:501-608).

* With W(omega_i) = p(omega_i^2) M on every mesh point, p a quadratic in
  omega^2 and M Hermitian, the 3-point omega^2 interpolation is exact, so
  zsec_mm' must equal the closed form sum over poles wfac * p(we^2) *
  z_m^H M z_m' -- where the poles are found by brute force over all
  (middle, external) pairs with the reference-built wfacx2/weavx2
  (oracle/_ref, S/wfacx.f90), not by the oracle's own window logic.
* Only the upper triangle of each W record is read (keepwv packing):
  changing the strict lower triangle leaves zsec unchanged.
* Core states (it <= nctot) take the unsmeared weight."""
import numpy as np
import pytest


def _mesh(nw):
    # ecalj's real mesh: freq_r(0) = 0, freq_r(i) = bin midpoints of the geometric histogram mesh
    aa, bb = 0.03, 1e-5 / 0.03
    fr = bb * (np.exp(aa * np.arange(nw + 2)) - 1.0)
    return np.concatenate([[0.0], 0.5 * (fr[:nw] + fr[1:nw + 1])])


def _setup(seed, ngb=9, nm=14, ntq=6):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((ngb, nm, ntq)) + 1j * rng.standard_normal((ngb, nm, ntq))
    ekc = np.sort(rng.uniform(-0.8, 0.6, nm))        # Ry
    omega = rng.uniform(-0.7, 0.5, ntq)              # Ry
    return rng, z, ekc, omega


def _wfac_we(ref, omg, ef, e, esmr):
    w = ref._QMm_wfacPwfacx2(C.c_double(omg), C.c_double(ef), C.c_double(e), C.c_double(esmr))
    return w


import ctypes as C  # noqa: E402


def _ref_fns():
    import oracle as orc

    ref = orc.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built here")
    f = ref._QMm_wfacPwfacx2
    g = ref._QMm_wfacPweavx2
    for fn in (f, g):
        fn.argtypes = [C.POINTER(C.c_double)] * 4
        fn.restype = C.c_double
    call = lambda fn, *a: fn(*[C.byref(C.c_double(x)) for x in a])
    return (lambda *a: call(f, *a)), (lambda *a: call(g, *a))


@pytest.mark.parametrize("seed,nctot", [(1, 0), (2, 3)])
def test_real_axis_polynomial_w_closed_form(orc, seed, nctot):
    wfacx2, weavx2 = _ref_fns()
    rng, z, ekc, omega = _setup(seed)
    ngb, nm, ntq = z.shape
    ef, esmr, wkkr = 0.02, 0.01, 0.125
    nw = 270
    fr = _mesh(nw)
    a = rng.standard_normal((ngb, ngb)) + 1j * rng.standard_normal((ngb, ngb))
    M = a + a.conj().T
    c0, c1, c2 = 0.7, -1.3, 0.4
    p = lambda x2: c0 + c1 * x2 + c2 * x2 * x2
    wr = np.stack([p(f * f) * M for f in fr], axis=2)
    zsec, ne = orc.sigc_real(z, wr, fr, omega, ekc, ef, esmr, wkkr, nctot=nctot)
    assert ne > 0
    ref = np.zeros((ntq, ntq), complex)
    npole = 0
    for itp in range(ntq):
        for it in range(nm):
            w = wfacx2(omega[itp], ef, ekc[it], 0.0 if it < nctot else esmr)
            if w < 1e-8:
                continue
            npole += 1
            w *= wkkr * (1.0 if omega[itp] - ef >= 0 else -1.0)
            we = 0.5 * abs(omega[itp] - weavx2(omega[itp], ef, ekc[it], esmr))
            v = w * p(we * we) * (M.conj().T @ z[:, it, itp])   # czmelwc column
            ref[itp, :] += v.conj() @ z[:, it, :]
    assert ne == 3 * npole
    for i in range(ntq):
        if ref[i, i].imag > 0:
            ref[i, i] = ref[i, i].real
    assert np.abs(zsec - ref).max() < 1e-11 * np.abs(ref).max()


def test_real_axis_reads_upper_triangle_only(orc):
    rng, z, ekc, omega = _setup(5)
    ngb = z.shape[0]
    nw = 270
    fr = _mesh(nw)
    wr = rng.standard_normal((ngb, ngb, nw + 1)) + 1j * rng.standard_normal((ngb, ngb, nw + 1))
    z1, _ = orc.sigc_real(z, wr, fr, omega, ekc, 0.0, 0.01, 0.2)
    w2 = wr.copy()
    low = np.tril_indices(ngb, -1)
    w2[low[0], low[1], :] = 123.0 - 7.0j
    z2, _ = orc.sigc_real(z, w2, fr, omega, ekc, 0.0, 0.01, 0.2)
    assert np.array_equal(z1, z2)


def test_real_axis_out_of_mesh_is_an_error(orc):
    rng, z, ekc, omega = _setup(6)
    fr = _mesh(20)   # far too short for the pole frequencies
    wr = np.zeros((z.shape[0], z.shape[0], 21), complex)
    with pytest.raises(ValueError):
        orc.sigc_real(z, wr, fr, omega, ekc, 0.0, 0.01, 0.2)


def test_full_equals_imag_plus_real_without_clipping(orc):
    """orc_sigc_full puts both parts in one czmelwc: with a Hermitian W on
    both axes (real diagonals) no clip acts, and it equals the sum of the
    separate imaginary-axis and real-axis results."""
    rng, z, ekc, omega = _setup(9)
    ngb, nm, ntq = z.shape
    niw, nw = 5, 270
    fr = _mesh(nw)
    herm = lambda a: a + a.conj().transpose(1, 0, 2)
    wc = herm(rng.standard_normal((ngb, ngb, niw + 1)) + 1j * rng.standard_normal((ngb, ngb, niw + 1)))
    wr = herm(rng.standard_normal((ngb, ngb, nw + 1)) + 1j * rng.standard_normal((ngb, ngb, nw + 1)))
    g = rng.standard_normal((niw + 1, nm, ntq))
    full, ne = orc.sigc_full(z, wc, g, wr, fr, omega, ekc, 0.0, 0.01, 0.3)
    a = orc.sigc_imag(z, wc, g)
    b, ne2 = orc.sigc_real(z, wr, fr, omega, ekc, 0.0, 0.01, 0.3)
    assert ne == ne2 > 0
    assert np.abs(full - (a + b)).max() < 1e-12 * np.abs(full).max()
