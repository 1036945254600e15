"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP library in
``ecalj_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package.  Every wrapper
takes and returns numpy arrays in Fortran (column-major) order; complex arrays
are complex128.  Each C function cites the ecalj routine it restates
(see oracle/oracle.h).

``ref()`` loads ``oracle/_ref/libecaljref.so``: ecalj's own
S/hilbertmat.f90 and S/wfacx.f90 compiled unmodified with flang by
oracle/Makefile, used to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build(force: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the ecalj tree is present)."""
    so = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        _LIB = C.CDLL(so)
        _LIB.orc_gauss.restype = C.c_int
        _LIB.orc_freq01.restype = C.c_int
        _LIB.orc_frhis.restype = C.c_int
        _LIB.orc_zmel.restype = C.c_int
        _LIB.orc_wc_build.restype = C.c_int
        _LIB.orc_zinv.restype = C.c_int
        _LIB.orc_wfacx2.restype = C.c_double
        _LIB.orc_weavx2.restype = C.c_double
        for name in ("orc_wfacx2", "orc_weavx2"):
            getattr(_LIB, name).argtypes = [C.c_double] * 4
        _LIB.orc_use_blas.argtypes = [C.c_void_p] * 3
        _LIB.orc_blas_active.restype = C.c_int
        use_blas(os.environ.get("ECALJ_ORACLE_LOOPS", "0") != "1")
    return _LIB


def _capsule_ptr(mod, name):
    cap = mod.__pyx_capi__[name]
    get_name = C.pythonapi.PyCapsule_GetName
    get_name.restype, get_name.argtypes = C.c_char_p, [C.py_object]
    get_ptr = C.pythonapi.PyCapsule_GetPointer
    get_ptr.restype, get_ptr.argtypes = C.c_void_p, [C.py_object, C.c_char_p]
    return get_ptr(cap, get_name(cap))


def use_blas(on: bool = True) -> bool:
    """Route the oracle's complex GEMMs and LU inverses through the LP64
    zgemm / zgetrf / zgetri that scipy ships (the routines the reference calls,
    S/m_blas.f90:225, S/m_lapack.f90:21-47), or (on=False) through the oracle's
    own loops.  Returns whether the BLAS path is active."""
    lib_ = lib()
    if not on:
        lib_.orc_use_blas(None, None, None)
        return False
    try:
        import scipy.linalg.cython_blas as cb
        import scipy.linalg.cython_lapack as cl
        lib_.orc_use_blas(_capsule_ptr(cb, "zgemm"), _capsule_ptr(cl, "zgetrf"), _capsule_ptr(cl, "zgetri"))
    except (ImportError, KeyError, AttributeError):
        lib_.orc_use_blas(None, None, None)
    return bool(lib_.orc_blas_active())


def ref():
    """The reference-compiled checker, or None if it was not built here."""
    global _REF
    if _REF is None:
        so = os.path.join(_HERE, "_ref", "libecaljref.so")
        if not os.path.exists(so):
            return None
        _REF = C.CDLL(so)
        _REF._QMm_wfacPwfacx2.restype = C.c_double
        _REF._QMm_wfacPweavx2.restype = C.c_double
    return _REF


def _p(a: np.ndarray):
    assert a.flags.f_contiguous or a.ndim <= 1, "oracle arrays must be Fortran-ordered"
    return a.ctypes.data_as(_dp)


def _ip_(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def _cz(x) -> np.ndarray:
    return np.asfortranarray(np.asarray(x, dtype=np.complex128))


def _cd(x) -> np.ndarray:
    return np.asfortranarray(np.asarray(x, dtype=np.float64))


def _ci(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


# --------------------------------------------------------------------------
def zgemm(opa, opb, a, b, alpha=1.0, beta=0.0, c=None):
    a = _cz(a)
    b = _cz(b)
    m = a.shape[0] if opa == "N" else a.shape[1]
    k = a.shape[1] if opa == "N" else a.shape[0]
    n = b.shape[1] if opb == "N" else b.shape[0]
    c = np.zeros((m, n), np.complex128, order="F") if c is None else _cz(c).copy(order="F")
    _ensure_cplx_sigs()
    lib().orc_zgemm(opa.encode(), opb.encode(), m, n, k, _cplx(alpha), _p(a), a.shape[0], _p(b),
                    b.shape[0], _cplx(beta), _p(c), m)
    return c


class _Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _cplx(z):
    z = complex(z)
    return _Cplx(z.real, z.imag)


lib_cplx_set = False


def _ensure_cplx_sigs():
    global lib_cplx_set
    if lib_cplx_set:
        return
    L = lib()
    L.orc_zgemm.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, _Cplx, _dp, C.c_long, _dp,
                            C.c_long, _Cplx, _dp, C.c_long]
    L.orc_hilbertmat.argtypes = [_Cplx, C.c_int, _dp, _dp, _dp, _dp]
    lib_cplx_set = True


def gauss(n, a=0.0, b=1.0):
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().orc_gauss(n, C.c_double(a), C.c_double(b), _p(x), _p(w)) != 0:
        raise ValueError(f"gauss: unsupported n={n}")
    return x, w


def freq01(nx, ua=1.0):
    fx, fw, wx, ex = (np.zeros(nx) for _ in range(4))
    if lib().orc_freq01(nx, C.c_double(ua), _p(fx), _p(fw), _p(wx), _p(ex)) != 0:
        raise ValueError(f"freq01: unsupported nx={nx}")
    return fx, fw, wx, ex


def frhis(omg2max, oratio=1.03, dw=1e-5):
    n = lib().orc_frhis(C.c_double(oratio), C.c_double(dw), C.c_double(omg2max), None)
    f = np.zeros(n + 1)
    lib().orc_frhis(C.c_double(oratio), C.c_double(dw), C.c_double(omg2max), _p(f))
    return n, f


def hilbertmat(zz, his_L, his_C, his_R):
    _ensure_cplx_sigs()
    nwhis = (len(his_L) - 1) // 2
    r = np.zeros(2 * nwhis + 1, np.complex128)
    lib().orc_hilbertmat(_cplx(zz), nwhis, _p(_cd(his_L)), _p(_cd(his_C)), _p(_cd(his_R)), _p(r))
    return r


def imatt(freqi, frhis_):
    freqi = _cd(freqi)
    frhis_ = _cd(frhis_)
    niw, nwhis = len(freqi), len(frhis_) - 1
    out = np.zeros((niw, nwhis), order="F")
    lib().orc_imatt(niw, _p(freqi), nwhis, _p(frhis_), _p(out))
    return out


def lindtet6(kvec, ea, eb, x, efa, efb, frhis_):
    """Histogram of one tetrahedron (kvec: 4 x 3 corners)."""
    fr = _cd(frhis_)
    nwhis = len(fr) - 1
    out = np.zeros(nwhis)
    k = np.ascontiguousarray(np.asarray(kvec, np.float64))
    rc = lib().orc_lindtet6(k.ctypes.data_as(_dp), _p(_cd(ea)), _p(_cd(eb)), _p(_cd(x)), C.c_double(efa),
                            C.c_double(efb), _p(fr), nwhis, _p(out))
    if rc != 0:
        raise ValueError(f"lindtet6 rc={rc}")
    return out


def lindtet6_ml(kvec, ea, eb, x, efa, efb, frhis_):
    """The same with tetrahedron_matrix_linear: nwhis x 4, column c the
    weight carried by corner c of kvec."""
    fr = _cd(frhis_)
    nwhis = len(fr) - 1
    out = np.zeros((nwhis, 4), order="F")
    k = np.ascontiguousarray(np.asarray(kvec, np.float64))
    rc = lib().orc_lindtet6_ml(k.ctypes.data_as(_dp), _p(_cd(ea)), _p(_cd(eb)), _p(_cd(x)), C.c_double(efa),
                               C.c_double(efb), _p(fr), nwhis, _p(out))
    if rc != 0:
        raise ValueError(f"lindtet6_ml rc={rc}")
    return np.ascontiguousarray(out)


def chkdgn(ene):
    e = _cd(ene)
    a = np.zeros(len(e), np.int32)
    b = np.zeros(len(e), np.int32)
    n = lib().orc_chkdgn(_p(e), len(e), _ip_(a), _ip_(b))
    return a[:n], b[:n]


def tetwt(mesh, ek1, ek2, ef, frhis_, ebmx=1e10, npm=1, ecore=None, matrix_linear=False):
    """gettetwt (npm = 1 or 2, core states of energies ecore appended after
    the nband valence bands, tetrahedron_matrix_linear): the reference's
    index sets (1-based values) and whw.  mesh: synth.TetraMesh.  With npm = 1 the arrays are (nbnbx, nqbz)
    and nbnb (nqbz,); with npm = 2 a trailing jpm axis is kept."""
    ek1 = _cd(ek1)
    ek2 = _cd(ek2)
    nband, nqbz = ek1.shape
    fr = _cd(frhis_)
    nwhis = len(fr) - 1
    ec = np.ascontiguousarray(np.zeros(0) if ecore is None else ecore, np.float64)
    nctot = len(ec)
    idt = np.asfortranarray(mesh.idtetf, np.int32)
    qw = np.asfortranarray(mesh.qbzw, np.float64)
    ib1 = np.ascontiguousarray(mesh.ib1bz, np.int32)
    ql = np.asfortranarray(mesh.qlat, np.float64)
    cap_b, cap_h = 1, 1
    for _ in range(3):  # sizes: nbnbx first, then nhwtot
        nbnb = np.zeros((nqbz, npm), np.int32, order="F")
        arrs = [np.zeros((cap_b, nqbz, npm), np.int32, order="F") for _ in range(5)]
        whw = np.zeros(cap_h)
        nbx, nht = C.c_int(0), C.c_int(0)
        rc = lib().orc_tetwt_full(nband, nqbz, idt.shape[1], _ip_(idt), _p(qw), _ip_(ib1), _p(ek1), _p(ek2),
                                  C.c_double(ef), C.c_double(ebmx), _p(ql), _p(fr), nwhis, npm, nctot,
                                  _p(ec) if nctot else None, int(bool(matrix_linear)), cap_b, cap_h, _ip_(nbnb), *[_ip_(a) for a in arrs],
                                  _p(whw), C.byref(nbx), C.byref(nht))
        if rc == -1:
            cap_b, cap_h = max(nbx.value, 1), max(nht.value, 1)
            continue
        if rc != 0:
            raise ValueError(f"orc_tetwt_full rc={rc}")
        if npm == 1:
            nbnb, arrs = nbnb[:, 0], [a[:, :, 0] for a in arrs]
        n1b, n2b, ihw, nhw, jhw = arrs
        return dict(nbnb=nbnb, n1b=n1b, n2b=n2b, ihw=ihw, nhw=nhw, jhw=jhw, whw=whw[:nht.value],
                    nbnbx=nbx.value, nhwtot=nht.value)
    raise RuntimeError("orc_tetwt_full: capacity loop did not converge")


def zmel_ipw(case, branch=0):
    """IPW block of zmel for a synth.IPWCase: ngc x nm x ntp0."""
    c = case
    ngc = c.ngveccR.shape[1]
    nm, ntp0 = c.dgeigqk.shape[1], c.geigq.shape[1]
    out = np.zeros((ngc, nm, ntp0), np.complex128, order="F")
    i32 = lambda a: np.asfortranarray(a, np.int32)
    args = [i32(c.ngveccR), i32(c.nadd), _cz(c.phase), i32(c.ngvecpB1), _cz(c.geigq), i32(c.ngvecpB2),
            _cz(c.dgeigqk), _cz(c.ggg), i32(c.ggg_lo), i32(np.array(c.igggi.shape)), i32(c.igggi),
            i32(c.nvgcgp2), i32(c.gcp_lo), i32(np.array(c.igcgp2i.shape)), i32(c.igcgp2i)]
    gcR, nadd, ph, g1, geq, g2, dgk, ggg, glo, gn, igg, nvg, clo, cn, icg = args
    rc = lib().orc_zmel_ipw(ngc, _ip_(gcR), _ip_(nadd), _p(ph), c.geigq.shape[0], g1.shape[1], _ip_(g1), _p(geq),
                            ntp0, g2.shape[1], _ip_(g2), _p(dgk), nm, _p(ggg), _ip_(glo), _ip_(gn), _ip_(igg),
                            nvg.shape[1], _ip_(nvg), _ip_(clo), _ip_(cn), _ip_(icg), branch, _p(out),
                            C.c_long(ngc))
    if rc != 0:
        raise ValueError(f"orc_zmel_ipw rc={rc}")
    return out


def rmatt(freqr, frhis_):
    """dpsion_init real branch (npm=1): (nw+1) x nwhis, freqr = freq_r(0:nw)."""
    freqr = _cd(freqr)
    frhis_ = _cd(frhis_)
    nw, nwhis = len(freqr) - 1, len(frhis_) - 1
    out = np.zeros((nw + 1, nwhis), order="F")
    lib().orc_rmatt(nw, _p(freqr), nwhis, _p(frhis_), _p(out))
    return out


def zmel(nv, mdim, phase, ppb, cphim, cphiq, m2e=None, conjg=False):
    nv = _ci(nv)
    mdim = _ci(mdim)
    phase = np.ascontiguousarray(np.asarray(phase, np.complex128))
    ppb = np.ascontiguousarray(np.asarray(ppb, np.float64).ravel(order="F"))
    cphim = _cz(cphim)
    cphiq = _cz(cphiq)
    ndima = cphim.shape[0]
    nm, nq = cphim.shape[1], cphiq.shape[1]
    ngb = int(mdim.sum())
    if m2e is not None:
        m2e = _cz(m2e)
        nbb = m2e.shape[1]
    else:
        nbb = ngb
    out = np.zeros((nbb, nm, nq), np.complex128, order="F")
    rc = lib().orc_zmel(len(nv), _ip_(nv), _ip_(mdim), _p(phase), _p(ppb), ndima, nm, _p(cphim), nq,
                        _p(cphiq), _p(m2e) if m2e is not None else None, nbb, int(bool(conjg)), _p(out))
    if rc != 0:
        raise ValueError(f"orc_zmel rc={rc}")
    return out


def x0gemm(rcxq, npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, zmel_, icoun_lo, icoun_hi, iwini, iwend,
           itc, itpc, jpmc, icouini, whwc):
    """Accumulates into rcxq (Fortran array npr x npr_col x nslice) in place."""
    assert rcxq.dtype == np.complex128 and rcxq.flags.f_contiguous
    zmel_ = _cz(zmel_)
    nbb, _, nqtot = zmel_.shape
    lib().orc_x0gemm(_p(rcxq), npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, _p(zmel_), nbb, nqtot,
                     icoun_lo, icoun_hi, _ip_(_ci(iwini)), _ip_(_ci(iwend)), _ip_(_ci(itc)),
                     _ip_(_ci(itpc)), _ip_(_ci(jpmc)), _ip_(_ci(icouini)), _p(_cd(whwc)))
    return rcxq


def dpsion_imag(rcxq, frhis_, imatt_, ecut=1e10):
    """Returns zxqi; rcxq (npr x npr_col x nwhis+1) is scaled in place as the reference does."""
    assert rcxq.dtype == np.complex128 and rcxq.flags.f_contiguous
    npr, npr_col, ns = rcxq.shape
    nwhis = ns - 1
    niw = imatt_.shape[0]
    zxqi = np.zeros((npr, npr_col, niw), np.complex128, order="F")
    lib().orc_dpsion_imag(npr, npr_col, nwhis, niw, _p(_cd(frhis_)), _p(_cd(imatt_)), C.c_double(ecut),
                          _p(rcxq), _p(zxqi))
    return zxqi


def chi0_freq(z, w):
    z = _cz(z)
    w = _cd(w)
    ni, npairs = z.shape
    nf = w.shape[0]
    chi = np.zeros((ni, ni, nf), np.complex128, order="F")
    lib().orc_chi0_freq(ni, npairs, nf, _p(z), _p(w), _p(chi))
    return chi


def dpsion_real(rcxq, frhis_, rmatt_, ecut=1e10, scaled=False):
    """Real-axis branch of dpsion_chiq (npm=1).  rcxq (npr x npr_col x nwhis+1,
    Fortran order) is scaled in place unless `scaled`; returns zxq(:,:,0:nw)."""
    assert rcxq.dtype == np.complex128 and rcxq.flags.f_contiguous
    npr, npr_col, ns = rcxq.shape
    nwhis = ns - 1
    nw = rmatt_.shape[0] - 1
    zxq = np.zeros((npr, npr_col, nw + 1), np.complex128, order="F")
    lib().orc_dpsion_real(npr, npr_col, nwhis, nw, _p(_cd(frhis_)), _p(_cd(rmatt_)), C.c_double(ecut),
                          int(bool(scaled)), _p(rcxq), _p(zxq))
    return zxq


def wc_build(chi, vcousq, m2e, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True):
    chi = _cz(chi)
    ngb, _, nw = chi.shape
    wc = np.zeros((ngb, ngb, nw), np.complex128, order="F")
    rc = lib().orc_wc_build(ngb, nw, nspin, int(is_x0_m_basis), int(is_wc_m_basis), _p(chi),
                            _p(_cd(vcousq)), _p(_cz(m2e)), _p(wc), ngb)
    if rc != 0:
        raise np.linalg.LinAlgError("orc_wc_build: singular epsilon")
    return wc


def zinv(a):
    a = _cz(a).copy(order="F")
    if lib().orc_zinv(a.shape[0], _p(a), C.c_long(a.shape[0])) != 0:
        raise np.linalg.LinAlgError("singular")
    return a


def wgtim(omega, ekc, esmr, wkkr, x, wt, expa, ua=1.0, ns1=1, nctot=0):
    omega = _cd(omega)
    ekc = _cd(ekc)
    niw = len(x)
    nm = len(ekc)
    ntq = len(omega)
    out = np.zeros((niw + 1, nm, ntq), order="F")
    lib().orc_wgtim(niw, ns1, ns1 + nm - 1, ntq, nctot, _p(omega), _p(ekc), C.c_double(esmr),
                    C.c_double(ua), C.c_double(wkkr), _p(_cd(x)), _p(_cd(wt)), _p(_cd(expa)), _p(out))
    return out


def sigc_imag(zmel_, wc, wgt, zsec=None):
    zmel_ = _cz(zmel_)
    wc = _cz(wc)
    ngb, nm, ntq = zmel_.shape
    niw = wc.shape[2] - 1
    zsec = np.zeros((ntq, ntq), np.complex128, order="F") if zsec is None else _cz(zsec).copy(order="F")
    lib().orc_sigc_imag(ngb, nm, ntq, niw, _p(zmel_), _p(wc), wc.shape[0], _p(_cd(wgt)), _p(zsec),
                        zsec.shape[0])
    return zsec


def wfacx2(e1, e2, ek, esmr):
    return lib().orc_wfacx2(e1, e2, ek, esmr)


def weavx2(e1, e2, ek, esmr):
    return lib().orc_weavx2(e1, e2, ek, esmr)


def qp_energy(elda, sex, sexcore, sec, vxc):
    sec = _cz(sec)
    ntq = len(elda)
    eqp = np.zeros(ntq)
    lib().orc_qp_energy(ntq, _p(_cd(elda)), _p(_cd(sex)), _p(_cd(sexcore)), _p(sec), sec.shape[0],
                        _p(_cd(vxc)), _p(eqp))
    return eqp


def sigma_herm(sex2, sexcore2, sec, vxcmat):
    ntq = sec.shape[0]
    se = np.zeros((ntq, ntq), np.complex128, order="F")
    lib().orc_sigma_herm(ntq, _p(_cz(sex2)), _p(_cz(sexcore2)), _p(_cz(sec)), _p(_cz(vxcmat)), _p(se))
    return se


# ---------------------------------------------------------------- learned-tail MLP (oracle/mlp.c)
def _pr(a):
    """Pointer to a C-ordered (row-major) float64 array; None passes NULL."""
    if a is None:
        return None
    assert a.flags.c_contiguous and a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def _mlp_sigs():
    L = lib()
    if getattr(L, "_mlp_ready", False):
        return L
    L.orc_mlp_nparam.restype = C.c_int64
    L.orc_mlp_nparam.argtypes = [C.c_int, _ip]
    L.orc_mlp_forward.argtypes = [C.c_int, _ip, _dp, _dp, C.c_int, _dp]
    L.orc_mlp_loss_grad.restype = C.c_double
    L.orc_mlp_loss_grad.argtypes = [C.c_int, _ip, _dp, _dp, _dp, C.c_int, _dp]
    for n in ("orc_silu", "orc_dsilu"):
        getattr(L, n).restype = C.c_double
        getattr(L, n).argtypes = [C.c_double]
    L._mlp_ready = True
    return L


def mlp_nparam(dims):
    d = _ci(dims)
    return int(_mlp_sigs().orc_mlp_nparam(len(d) - 1, _ip_(d)))


def mlp_forward(dims, params, x):
    """y = MLP(x); x: B x dims[0] (row-major), params float64 (nparam)."""
    d = _ci(dims)
    x = np.ascontiguousarray(x, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    y = np.zeros((x.shape[0], d[-1]))
    _mlp_sigs().orc_mlp_forward(len(d) - 1, _ip_(d), _pr(p), _pr(x), x.shape[0], _pr(y))
    return y


def mlp_loss_grad(dims, params, x, t, want_grad=True):
    """(mean squared error, its gradient w.r.t. params or None)."""
    d = _ci(dims)
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    p = np.ascontiguousarray(params, np.float64)
    g = np.zeros(p.size) if want_grad else None
    loss = _mlp_sigs().orc_mlp_loss_grad(len(d) - 1, _ip_(d), _pr(p), _pr(x), _pr(t), x.shape[0],
                                         _pr(g) if want_grad else None)
    return float(loss), g


def silu(z):
    return _mlp_sigs().orc_silu(float(z))


def dsilu(z):
    return _mlp_sigs().orc_dsilu(float(z))


def sigc_real(zmel_, wr, freq_r, omega, ekc, ef, esmr, wkkr, nctot=0, ns2r=None, zsec=None):
    """Real-axis pole term of Sigma_c for one internal k (oracle.c orc_sigc_real).
    zmel: ngb x nm x ntq; wr: ngb x ngb x (nw+1); freq_r: nw+1 (Hartree); energies Ry.
    Returns (zsec, number of entries)."""
    z = _cz(zmel_)
    w = _cz(wr)
    ngb, nm, ntq = z.shape
    nw = w.shape[2] - 1
    fr = _cd(freq_r)
    assert fr.size == nw + 1
    if zsec is None:
        zsec = np.zeros((ntq, ntq), np.complex128, order="F")
    L = lib()
    L.orc_sigc_real.restype = C.c_int
    L.orc_sigc_real.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp, C.c_double,
                                C.c_double, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
    ne = L.orc_sigc_real(ngb, nm, ntq, _p(z), _p(w), w.shape[0], nw, _p(fr), _p(_cd(omega)), _p(_cd(ekc)), ef, esmr,
                         nctot, nm if ns2r is None else ns2r, wkkr, _p(zsec), zsec.shape[0])
    if ne < 0:
        raise ValueError("sigc_real: a pole frequency lies outside the real mesh")
    return zsec, ne


def sigc_full(zmel_, wc, wgt, wr, freq_r, omega, ekc, ef, esmr, wkkr, nctot=0, ns2r=None, zsec=None):
    """Imaginary + real-axis Sigma_c of one internal k in one czmelwc (orc_sigc_full)."""
    z, w, g, r = _cz(zmel_), _cz(wc), _cd(wgt), _cz(wr)
    ngb, nm, ntq = z.shape
    niw = w.shape[2] - 1
    nw = r.shape[2] - 1
    if zsec is None:
        zsec = np.zeros((ntq, ntq), np.complex128, order="F")
    L = lib()
    L.orc_sigc_full.restype = C.c_int
    L.orc_sigc_full.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double, _dp, C.c_int]
    ne = L.orc_sigc_full(ngb, nm, ntq, niw, _p(z), _p(w), w.shape[0], _p(g), _p(r), r.shape[0], nw, _p(_cd(freq_r)),
                         _p(_cd(omega)), _p(_cd(ekc)), ef, esmr, nctot, nm if ns2r is None else ns2r, wkkr, _p(zsec),
                         zsec.shape[0])
    if ne < 0:
        raise ValueError("sigc_full: a pole frequency lies outside the real mesh")
    return zsec, ne


# --------------------------------------------------------------------------
# Gamma point (oracle/gamma.c)
def sylm(r, lmx):
    """sylm * sylmnc: normalized real harmonics at the unit vector of r."""
    r = _cd(r)
    n = (lmx + 1) ** 2
    yl, cy = np.zeros(n), np.zeros(n)
    lib().orc_sylm(_p(r), _p(yl), lmx)
    lib().orc_sylmnc(_p(cy), lmx)
    return yl * cy / np.linalg.norm(r) ** np.array([int(np.sqrt(i)) for i in range(n)])


def gaunt(l12, l3):
    """g[L1, L2, L3] = int Y_L1 Y_L2 Y_L3 (0-based L)."""
    n12, n3 = (l12 + 1) ** 2, (l3 + 1) ** 2
    g = np.zeros((n12, n12, n3), order="F")
    if lib().orc_gaunt(l12, l3, _p(g)) != 0:
        raise ValueError("orc_gaunt: l out of range")
    return g


def getw0(ii, llw, dmlx, epinvq0i, wklm, wbz, lmxax, epinv):
    """getw0: llw (nw, nq0i) complex for frequencies ii..ii+nw-1."""
    llw = _cz(llw)
    nw, nq0i = llw.shape
    w0 = np.zeros(nw, np.complex128)
    llmat = np.zeros((3, 3), np.complex128, order="F")
    rc = lib().orc_getw0(ii, ii + nw - 1, nq0i, lmxax, _p(llw), _p(_cd(dmlx)), _p(_cd(epinvq0i)), _p(_cd(wklm)),
                         C.c_double(wbz), _p(_cd(epinv)), _p(w0), _p(llmat))
    if rc != 0:
        raise np.linalg.LinAlgError("orc_getw0: singular")
    return w0, llmat


def wc_build_gamma(chi, vcousq, m2e, nspin=1, is_x0_m_basis=True):
    chi = _cz(chi)
    ngb, _, nw = chi.shape
    wc = np.zeros((ngb, ngb, nw), np.complex128, order="F")
    rc = lib().orc_wc_build_gamma(ngb, nw, nspin, int(is_x0_m_basis), _p(chi), _p(_cd(vcousq)), _p(_cz(m2e)),
                                  _p(wc), ngb)
    if rc != 0:
        raise np.linalg.LinAlgError("orc_wc_build_gamma: singular epsilon")
    return wc


def eps_head(chi, vcousq, m2e, vcou1, nspin=1, is_x0_m_basis=True, wing_iw=-1):
    chi = _cz(chi)
    ngb, _, nw = chi.shape
    llw = np.zeros(nw, np.complex128)
    wing = np.zeros(max(ngb - 1, 1), np.complex128)
    rc = lib().orc_eps_head(ngb, nw, nspin, int(is_x0_m_basis), _p(chi), _p(_cd(vcousq)), _p(_cz(m2e)),
                            C.c_double(vcou1), _p(llw), wing_iw, _p(wing))
    if rc != 0:
        raise np.linalg.LinAlgError("orc_eps_head: singular epsilon")
    return llw, wing[:ngb - 1]


def wv0_patch(w, w0, m2e, is_wc_m_basis=True):
    w = _cz(w).copy(order="F")
    ngb, _, nw = w.shape
    lib().orc_wv0_patch(ngb, nw, _p(w), ngb, _p(_cz(w0)), int(is_wc_m_basis), _p(_cz(m2e)))
    return w


# --------------------------------------------------------------------------
# Eigenfunction rotation (oracle/rotwave.c)
def rotmto(case, iq, cphi_in):
    """rotmto for target iq (1-based into case.qtt): cphi_in ndima x nband at qtt(:, iqmap(iq))."""
    c = case
    iqq, igg = c.iqmap[iq - 1], c.igmap[iq - 1]
    cin = _cz(cphi_in)
    out = np.zeros_like(cin, order="F")
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    rc = lib().orc_rotmto(_p(_cd(c.qtt[:, iqq - 1])), cin.shape[0], cin.shape[1], len(c.ibas_tbl), _ip_(i32(c.ibas_tbl)),
                          _ip_(i32(c.l_tbl)), _ip_(i32(c.k_tbl)), _ip_(i32(c.offset_tbl)),
                          _p(_cd(c.symops[:, :, igg - 1])), _p(_cd(c.dlmm[:, :, :, igg - 1])), c.lmxax,
                          _ip_(i32(c.miat[:, igg - 1])), _p(_cd(c.tiat[:, :, igg - 1])), c.nbas, _p(cin), _p(out))
    if rc != 0:
        raise ValueError("orc_rotmto: orbital tables do not map a block")
    return out


def rotipw(case, iq, geig_in):
    """rotipw for target iq: geig_in ngp(iqq) x nband at qtt(:, iqmap(iq))."""
    c = case
    iqq, igg = c.iqmap[iq - 1], c.igmap[iq - 1]
    gin = _cz(geig_in)
    out = np.zeros_like(gin, order="F")
    platt = np.asfortranarray(c.plat.T)
    rc = lib().orc_rotipw(_p(_cd(c.qtt[:, iqq - 1])), _p(_cd(c.qtt[:, iq - 1])), int(c.ngp[iqq - 1]), gin.shape[1],
                          _p(platt), _p(_cd(c.qlat)), _p(_cd(c.symops[:, :, igg - 1])),
                          _ip_(np.asfortranarray(c.ngvecp[:, :, iqq - 1], np.int32)),
                          _ip_(np.asfortranarray(c.ngvecprev[:, :, :, iq - 1], np.int32)), c.imx,
                          _p(_cd(c.shtvg[:, igg - 1])), _p(gin), _p(out), C.c_long(gin.shape[0]))
    if rc != 0:
        raise ValueError("orc_rotipw: rotated G outside the target list")
    return out


def x0kf_init(tables, whw, nband, nctot=0, nwhis=None):
    """x0kf_v4hz_init (S/x0kf_v4h.f90:48-142) on tetwt's per-k pair tables.

    tables: one entry per jpm (npm = len(tables)); each entry a tuple of
    per-k lists (n1b, n2b, ihw, nhw, jhw) of 1-based int arrays, as gettetwt
    returns them; whw: all weights (jhw points into it, 1-based).  Returns a
    namespace with the fields of the index lists (kc, itc, itpc, jpmc, iwini,
    iwend, icouini, whwc, icounkmin, icounkmax, nkmin, nkmax, nkqmin, nkqmax)."""
    from types import SimpleNamespace

    npm = len(tables)
    nqbz = len(tables[0][0])
    nbnb = np.zeros((nqbz, npm), np.int32, order="F")
    for j in range(npm):
        for k in range(nqbz):
            nbnb[k, j] = len(tables[j][0][k])
    nbnbx = max(1, int(nbnb.max()))
    arrs = []
    for f in range(5):
        a = np.zeros((nbnbx, nqbz, npm), np.int32, order="F")
        for j in range(npm):
            for k in range(nqbz):
                v = np.asarray(tables[j][f][k], np.int32)
                a[:len(v), k, j] = v
        arrs.append(a)
    if nwhis is None:
        nwhis = max(int((arrs[2] + arrs[3] - 1).max()), 1)
    cap_coun = max(1, int(nbnb.sum()))
    cap_count = max(1, int(sum(int(np.sum(tables[j][3][k])) for j in range(npm) for k in range(nqbz))))
    o = {n: np.zeros(cap_coun, np.int32) for n in ("kc", "itc", "itpc", "jpmc", "iwini", "iwend", "icouini")}
    whwc = np.zeros(cap_count)
    per_k = {n: np.zeros(nqbz, np.int32) for n in ("icounkmin", "icounkmax", "nkmin", "nkmax", "nkqmin", "nkqmax")}
    ncoun, ncount = C.c_long(0), C.c_long(0)
    whw = np.ascontiguousarray(whw, np.float64)
    r = lib().orc_x0kf_init(C.c_int(nqbz), C.c_int(npm), C.c_int(nbnbx), _p(nbnb), *[_p(a) for a in arrs], _p(whw),
                            C.c_int(nband), C.c_int(nctot), C.c_int(nwhis), C.c_long(cap_coun), C.c_long(cap_count),
                            *[_p(o[n]) for n in ("kc", "itc", "itpc", "jpmc", "iwini", "iwend", "icouini")],
                            _p(whwc), *[_p(per_k[n]) for n in ("icounkmin", "icounkmax", "nkmin", "nkmax", "nkqmin",
                                                               "nkqmax")],
                            C.byref(ncoun), C.byref(ncount))
    if r:
        raise ValueError(f"orc_x0kf_init: status {r}")
    n1, n2 = ncoun.value, ncount.value
    out = {k: v[:n1].copy() for k, v in o.items()}
    out["whwc"] = whwc[:n2].copy()
    out.update(per_k)
    return SimpleNamespace(**out)


def real_mesh(eq_ry, ebz_ry, esmr, frhis, deltaw=0.02, ffac=10.0, ef=0.0):
    """getwemax + getfreq's real mesh (S/m_freq.f90:197-268, 110-129): freq_r(0:nw), Hartree."""
    eq = np.ascontiguousarray(eq_ry, np.float64)
    ebz = np.ascontiguousarray(ebz_ry, np.float64).ravel()
    fr = np.ascontiguousarray(frhis, np.float64)
    nwhis = len(fr) - 1
    out = np.zeros(nwhis + 1)
    wemax = C.c_double(0.0)
    L = lib()
    L.orc_real_mesh.restype = C.c_int
    nw = L.orc_real_mesh(C.c_int(len(eq)), _p(eq), C.c_long(len(ebz)), _p(ebz), C.c_double(ef), C.c_double(deltaw),
                         C.c_double(esmr), C.c_double(ffac), C.c_int(nwhis), _p(fr), C.byref(wemax), _p(out),
                         C.c_int(len(out)))
    if nw < 0:
        raise ValueError("real_mesh: histogram mesh does not reach wemax")
    return out[:nw + 1].copy()


def zmel_init(nlnmc, nlnmv, nblocha, ppb_full, ns1, ns2, nctot, cphim, cphiq, phase=None, iatomp=None, m2e=None,
              conjg=False):
    """get_zmel_init_gemm's MT part with core states and iatomp (orc_zmel_init).
    ppb_full: (nlnmx, nlnmx, mdimx, natom) real; cphim: ndima x (valence bands in ns1..ns2)."""
    nlnmc, nlnmv, nblocha = _ci(nlnmc), _ci(nlnmv), _ci(nblocha)
    natom = len(nlnmc)
    ppb_full = np.asfortranarray(ppb_full, np.float64)
    nlnmx, _, mdimx, _ = ppb_full.shape
    phase = np.ascontiguousarray(np.ones(natom, np.complex128) if phase is None else np.asarray(phase, np.complex128))
    cphim = _cz(cphim)
    cphiq = _cz(cphiq)
    ndima, nq = cphiq.shape
    ngb = int(nblocha.sum())
    nbb = m2e.shape[1] if m2e is not None else ngb
    out = np.zeros((nbb, ns2 - ns1 + 1, nq), np.complex128, order="F")
    ia = _ci(iatomp) if iatomp is not None else None
    rc = lib().orc_zmel_init(natom, nlnmx, mdimx, _ip_(nlnmc), _ip_(nlnmv), _ip_(nblocha),
                             _ip_(ia) if ia is not None else None, _p(phase), _p(ppb_full), ndima, ns1, ns2, nctot,
                             _p(cphim) if cphim.size else _p(np.zeros(1, np.complex128)), nq, _p(cphiq),
                             _p(_cz(m2e)) if m2e is not None else None, nbb, int(bool(conjg)), _p(out))
    if rc != 0:
        raise ValueError(f"orc_zmel_init rc={rc}")
    return out
