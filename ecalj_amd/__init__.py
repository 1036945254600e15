"""ecalj_amd -- MI355X (gfx950) implementation of ecalj's QSGW chi0 -> W -> Sigma_c path.

The compute lives in ``libecalj_gpu.so`` (HIP kernels behind the C ABI of
``include/ecalj_gpu.h``).  This module is the host-side mirror of the
reference's interface for the path: a handle holding the module state that
ecalj keeps in Fortran modules (ppb of m_zmel, the Coulomb eigenbasis of
m_readVcoud) and one method per entry point, with the reference's argument
meaning.  Host methods take numpy arrays (Fortran order); ``*_dev`` methods
take ``DevArray`` buffers allocated through the library, so that the process
uses the library's HIP runtime only.  There is no CPU fallback: if the
library is missing or no GPU is present, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libecalj_gpu.so")
_LIB = None

PREC = {"f64": 0, "f32": 1, "bf16x3": 2, "bf16f8": 3}
KCLASS = {"sigc": 0, "chi0": 1, "zmel": 2, "wbuild": 3, "hilbert": 4}


class EcaljError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = lib().ecalj_status_string(status).decode()
        super().__init__(f"{where}: {msg} (status {status})")
        self.status = status


# (name, argtypes); every entry point returns int status
_i, _l, _d, _p = C.c_int, C.c_int64, C.c_double, C.c_void_p
_SIGS = {
    "ecalj_create": [C.POINTER(C.c_void_p), _i],
    "ecalj_destroy": [_p],
    "ecalj_set_precision": [_p, _i],
    "ecalj_set_stream": [_p, _p],
    "ecalj_sync": [_p],
    "ecalj_stream_wait": [_p, _p],
    "ecalj_kernel_timing": [_p, _i],
    "ecalj_kernel_time": [_p, _i, C.POINTER(_d), C.POINTER(_l)],
    "ecalj_set_tuning": [_p, _i],
    "ecalj_graph_begin": [_p],
    "ecalj_graph_end": [_p, C.POINTER(C.c_void_p)],
    "ecalj_graph_launch": [_p, _p],
    "ecalj_graph_destroy": [_p],
    "ecalj_tuning_report": [_p, C.c_char_p, _l],
    "ecalj_set_ppb": [_p, _i, _p, _p, _p],
    "ecalj_set_coulomb": [_p, _i, _p, _p],
    "ecalj_zhgv": [_p, _p, _p, _i, _p, _i, _i, _i, _i],
    "ecalj_zhgv_dev": [_p, _p, _p, _i, _p, _i, _i, _i, _i],
    "ecalj_zmel": [_p, _p, _i, _i, _p, _i, _p, _i, _i, _i, _p],
    "ecalj_zmel_dev": [_p, _p, _i, _i, _p, _i, _p, _i, _i, _i, _p],
    "ecalj_zmel_strided_dev": [_p, _p, _i, _i, _p, _l, _i, _p, _l, _i, _i, _i, _p],
    "ecalj_set_ppb_full": [_p, _i, _i, _i, _p, _p, _p, _p],
    "ecalj_zmel_init": [_p, _p, _p, _i, _i, _i, _i, _p, _i, _p, _i, _i, _i, _p],
    "ecalj_zmel_init_dev": [_p, _p, _p, _i, _i, _i, _i, _p, _i, _p, _i, _i, _i, _p],
    "ecalj_x0gemm": [_p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p, _p],
    "ecalj_x0gemm_batched_dev": [_p, _p, _i, _i, _i, _i, _i, _i, _p, _p, _p, _l, _i, _i, _p, _p, _i, _p, _p,
                                 _p, _p, _p, _p, _i, _p],
    "ecalj_dpsion_imag": [_p, _p, _i, _i, _i, _i, _p, _p, _d, _i, _p],
    "ecalj_dpsion_imag_dev": [_p, _p, _i, _i, _i, _i, _p, _p, _d, _i, _p],
    "ecalj_chi0_freq": [_p, _p, _p, _i, _i, _i, _p],
    "ecalj_chi0_freq_dev": [_p, _p, _p, _i, _i, _i, _p],
    "ecalj_wc_build": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "ecalj_wc_build_dev": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "ecalj_wc_build_real": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "ecalj_wc_build_real_dev": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "ecalj_dpsion_real": [_p, _p, _i, _i, _i, _i, _p, _p, _d, _i, _p],
    "ecalj_dpsion_real_dev": [_p, _p, _i, _i, _i, _i, _p, _p, _d, _i, _p],
    "ecalj_zminv": [_p, _p, _i, _i],
    "ecalj_vcoud_write": [C.c_char_p, _i, _p, _p, _p],
    "ecalj_vcoud_read": [C.c_char_p, _p, _p, _p, _p, _i],
    "ecalj_sec2_write": [C.c_char_p, _p, _i, _i, _p, _p, _p],
    "ecalj_sec2_read": [C.c_char_p, _p, _i, _i, _p, _p, _p],
    "ecalj_wv_d_write": [C.c_char_p, _p],
    "ecalj_wv_d_read": [C.c_char_p, _p],
    "ecalj_wv_write_dev": [_p, C.c_char_p, _i, _i, _i, _p],
    "ecalj_wv_read_dev": [_p, C.c_char_p, _i, _i, _i, _p],
    "ecalj_wv4_write_dev": [_p, C.c_char_p, _i, _i, _i, _p],
    "ecalj_wv4_read_dev": [_p, C.c_char_p, _i, _i, _i, _p],
    "ecalj_set_ppovl": [_p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _p],
    "ecalj_zmel_ipw": [_p, _i, _p, _p, _p, _i, _i, _p, _p, _i, _i, _p, _p, _i, _p, _i, _i],
    "ecalj_zmel_ipw_dev": [_p, _i, _p, _p, _p, _i, _i, _p, _p, _i, _i, _p, _p, _i, _p, _i, _i],
    "ecalj_zmel_assemble_dev": [_p, _p, _i, _p, _i, _l, _i, _i, _p],
    "ecalj_tetwt": [_p, _i, _i, _i, _p, _i, _p, _p, _p, _p, _d, _d, _p, _p, _i, _i, _l, _p, _p, _p, _p, _p, _p,
                    _p, _p, _p],
    "ecalj_tetwt_full": [_p, _i, _i, _i, _p, _i, _p, _p, _p, _p, _d, _d, _p, _p, _i, _i, _i, _p, _i, _i, _l, _p,
                         _p, _p, _p, _p, _p, _p, _p, _p],
    "ecalj_zminv_dev": [_p, _p, _i, _i, _i, C.c_int64],
    "ecalj_zminv_noalloc_dev": [_p, _p, _i, _i, _i, C.c_int64],
    "ecalj_wc_build_gamma": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "ecalj_wc_build_gamma_dev": [_p, _p, _i, _i, _i, _i, _i, _p, _i],
    "ecalj_eps_head": [_p, _p, _i, _i, _i, _i, _d, _i, _p, _i, _p],
    "ecalj_eps_head_dev": [_p, _p, _i, _i, _i, _i, _d, _i, _p, _i, _p],
    "ecalj_getw0": [_i, _i, _i, _i, _p, _p, _p, _p, _d, _p, _p, _p],
    "ecalj_wv0_patch": [_p, _p, _i, _i, _i, _p, _i],
    "ecalj_wv0_patch_dev": [_p, _p, _i, _i, _i, _p, _i],
    "ecalj_set_hamindex": [_p, _i, _p, _p, _p, _p, _p, _i, _i, _p, _p, _i, _p, _p, _i, _p, _p, _i, _p, _p, _p],
    "ecalj_set_mnla": [_p, _i, _p, _p, _p, _p],
    "ecalj_readcphi_dev": [_p, _i, _p, _i, _i, _p, _i, _p],
    "ecalj_readgeig_dev": [_p, _i, _p, _i, _i, _p, _i, _p],
    "ecalj_mnla_read": [C.c_char_p, _i, _p, _p, _p, _p, _p],
    "ecalj_qgpsi_read": [C.c_char_p, _p, _p, _i, _i, _i, _p, _p, _p, _p, _p],
    "ecalj_evalue_read": [C.c_char_p, _p, _l, _p, _p],
    "ecalj_rec_read_dev": [_p, C.c_char_p, _l, _i, _i, _i, _l, _p, _l],
    "ecalj_sigc_weights": [_p, _i, _i, _i, _i, _i, _p, _p, _d, _d, _d, _p, _p, _p, _p],
    "ecalj_sigc_weights_dev": [_p, _i, _i, _i, _i, _i, _i, _p, _p, _d, _d, _p, _p, _p, _p, _p],
    "ecalj_sigc": [_p, _p, _i, _i, _i, _i, _p, _i, _p, _p, _i],
    "ecalj_sigc_dev": [_p, _i, _p, _i, _i, _i, _i, _p, _i, _p, _p, _i],
    "ecalj_sigc_grouped_dev": [_p, _i, _p, _p, _i, _i, _i, _p, _i, _p, _p, _i],
    "ecalj_sigc_real_dev": [_p, _i, _p, _i, _i, _i, _p, _i, _l, _i, _p, _p, _p, _d, _d, _i, _i, _p, _p, _i],
    "ecalj_sigc_full_dev": [_p, _i, _p, _i, _i, _i, _i, _p, _i, _p, _p, _i, _l, _i, _p, _p, _p, _d, _d, _i, _i,
                            _p, _p, _i],
    "ecalj_sigc_real": [_p, _p, _i, _i, _i, _p, _i, _i, _p, _p, _p, _d, _d, _i, _i, _d, _p, _i],
    "ecalj_qp_energy": [_i, _p, _p, _p, _p, _i, _p, _p],
    "ecalj_dev_alloc": [_p, C.c_size_t, C.POINTER(C.c_void_p)],
    "ecalj_dev_free": [_p, _p],
    "ecalj_memcpy_h2d": [_p, _p, _p, C.c_size_t],
    "ecalj_memcpy_d2h": [_p, _p, _p, C.c_size_t],
    "ecalj_memcpy_d2d": [_p, _p, _p, C.c_size_t],
    "ecalj_dev_zero": [_p, _p, C.c_size_t],
    "ecalj_dev_copy": [_p, _p, _p, C.c_size_t],
    "ecalj_dev_phase_similarity": [_p, _p, _i, _i, _p, _i, _p],
    "ecalj_sigma_herm": [_i, _p, _p, _p, _p, _p],
    "ecalj_mlp_nparam": [_i, _p, C.POINTER(_l)],
    "ecalj_mlp_forward_dev": [_p, _i, _p, _p, _p, _i, _p],
    "ecalj_mlp_grad_dev": [_p, _i, _p, _p, _p, _p, _i, _p, _p],
    "ecalj_mlp_adam_dev": [_p, _l, _p, _p, _p, _p, _i, _d, _d, _d, _d],
    "ecalj_mlp_to_bf16_dev": [_p, _l, _p, _p, _p],
    "ecalj_mlp_infer_bf16_dev": [_p, _i, _p, _p, _p, _p, _p, _i, _p],
    "ecalj_tail_features_dev": [_p, _p, _i, _i, _i, _p, _p, _p, _p],
    "ecalj_tail_apply_dev": [_p, _p, _i, _i, _i, _p, _p, _p, _p, _p],
}


def lib():
    """Load libecalj_gpu.so; raises if it was not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run python -m ecalj_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.ecalj_status_string.argtypes = [C.c_int]
        L.ecalj_status_string.restype = C.c_char_p
        _LIB = L
    return _LIB


def getw0(ii, llw, dmlx, epinvq0i, wklm, wbz, lmxax, epinv):
    """getw0 (S/m_w0w0i.f90:146-311), host FP64: w0 for frequencies ii..ii+nw-1 and llmat (iw = 0)."""
    L = np.asfortranarray(llw, np.complex128)
    nw, nq0i = L.shape
    w0 = np.zeros(nw, np.complex128)
    llmat = np.zeros((3, 3), np.complex128, order="F")
    arrs = [np.asfortranarray(a, np.float64) for a in (dmlx, epinvq0i, wklm, epinv)]
    _chk(lib().ecalj_getw0(ii, ii + nw - 1, nq0i, lmxax, _np_ptr(L), *[_np_ptr(a) for a in arrs[:3]], float(wbz),
                           _np_ptr(arrs[3]), _np_ptr(w0), _np_ptr(llmat)), "getw0")
    return w0, llmat


def mnla_read(path):
    """@MNLA_CPHI tables: (ibas_tbl, l_tbl, k_tbl, offset_tbl)."""
    n = C.c_int(0)
    rc = lib().ecalj_mnla_read(str(path).encode(), 0, C.byref(n), None, None, None, None)
    if rc not in (0, 1):
        _chk(rc, "mnla_read")
    t = [np.zeros(n.value, np.int32) for _ in range(4)]
    _chk(lib().ecalj_mnla_read(str(path).encode(), n.value, C.byref(n), *[_np_ptr(a) for a in t]), "mnla_read")
    return tuple(t)


def qgpsi_read(path):
    """__QGpsi: dict(nqnum, ngpmx, nqbz, nqi, imx, nqibz, qpgcut, qtt, ngp, irr, ngvecp, ngvecprev)."""
    h6 = np.zeros(6, np.int32)
    cut = C.c_double(0.0)
    rc = lib().ecalj_qgpsi_read(str(path).encode(), _np_ptr(h6), C.byref(cut), 0, 0, 0, None, None, None, None, None)
    if rc not in (0, 1):
        _chk(rc, "qgpsi_read")
    nq, ngpmx, _, _, imx, _ = (int(x) for x in h6)
    nc = 2 * imx + 1
    qtt = np.zeros((3, nq), order="F")
    ngp, irr = np.zeros(nq, np.int32), np.zeros(nq, np.int32)
    gv = np.zeros((3, ngpmx, nq), np.int32, order="F")
    gr = np.zeros((nc, nc, nc, nq), np.int32, order="F")
    _chk(lib().ecalj_qgpsi_read(str(path).encode(), _np_ptr(h6), C.byref(cut), nq, ngpmx, imx, _np_ptr(qtt),
                                _np_ptr(ngp), _np_ptr(irr), _np_ptr(gv), _np_ptr(gr)), "qgpsi_read")
    keys = ("nqnum", "ngpmx", "nqbz", "nqi", "imx", "nqibz")
    return dict(zip(keys, (int(x) for x in h6)), qpgcut=cut.value, qtt=qtt, ngp=ngp, irr=irr, ngvecp=gv,
                ngvecprev=gr)


def evalue_read(path):
    """__EValue: dict(nband, nqi, nsp, nspc, nspx, qtti, evud)."""
    h5 = np.zeros(5, np.int32)
    rc = lib().ecalj_evalue_read(str(path).encode(), _np_ptr(h5), 0, None, None)
    if rc not in (0, 1):
        _chk(rc, "evalue_read")
    nband, nqi, nsp, nspc, nspx = (int(x) for x in h5)
    qtti = np.zeros((3, nqi), order="F")
    ev = np.zeros((nband, nqi, nspx), order="F")
    _chk(lib().ecalj_evalue_read(str(path).encode(), _np_ptr(h5), ev.size, _np_ptr(qtti), _np_ptr(ev)), "evalue_read")
    return dict(nband=nband, nqi=nqi, nsp=nsp, nspc=nspc, nspx=nspx, qtti=qtti, evud=ev)


def mlp_nparam(dims) -> int:
    d = _i32(dims)
    n = C.c_int64()
    _chk(lib().ecalj_mlp_nparam(len(d) - 1, _np_ptr(d), C.byref(n)), "mlp_nparam")
    return n.value


def exported_symbols():
    return list(_SIGS) + ["ecalj_status_string"]


def _chk(status, where):
    if status != 0:
        raise EcaljError(status, where)


def _np_ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _f(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _t_ptr(t):
    """Device pointer of a DevArray (or a raw integer address)."""
    if t is None:
        return None
    if isinstance(t, DevArray):
        return C.c_void_p(t.ptr)
    return C.c_void_p(int(t))


class DevArray:
    """A device buffer owned by a Handle; shape is Fortran-ordered."""

    def __init__(self, h: "Handle", shape, dtype):
        self.h = h
        self.shape = tuple(int(x) for x in (shape if np.iterable(shape) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        _chk(lib().ecalj_dev_alloc(h.ptr, self.nbytes, C.byref(p)), "dev_alloc")
        self.ptr = p.value
        h._arrays.add(self)  # freed by the handle's close() at the latest

    @classmethod
    def from_host(cls, h, a, dtype=None):
        a = np.asfortranarray(a if dtype is None else np.asarray(a, dtype))
        d = cls(h, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asfortranarray(np.asarray(a, self.dtype))
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _chk(lib().ecalj_memcpy_h2d(self.h.ptr, C.c_void_p(self.ptr), _np_ptr(a), self.nbytes), "memcpy_h2d")

    def download(self):
        out = np.empty(self.shape, self.dtype, order="F")
        _chk(lib().ecalj_memcpy_d2h(self.h.ptr, _np_ptr(out), C.c_void_p(self.ptr), self.nbytes), "memcpy_d2h")
        return out

    def zero(self):
        _chk(lib().ecalj_dev_zero(self.h.ptr, C.c_void_p(self.ptr), self.nbytes), "dev_zero")

    def free(self):
        if self.ptr:
            lib().ecalj_dev_free(self.h.ptr, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.h.ptr:
                self.free()
        except Exception:
            pass


class Graph:
    """A captured HIP graph of a handle's stream (ecalj_graph_*)."""

    def __init__(self, h: "Handle", ptr):
        self.h, self.ptr = h, ptr

    def launch(self):
        _chk(lib().ecalj_graph_launch(self.h.ptr, C.c_void_p(self.ptr)), "graph_launch")

    def close(self):
        if self.ptr:
            lib().ecalj_graph_destroy(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Handle:
    """One library handle (device binding, stream, module state, scratch)."""

    def __init__(self, device: int = 0, precision: str = "f64"):
        self._h = C.c_void_p()
        self._arrays = weakref.WeakSet()
        _chk(lib().ecalj_create(C.byref(self._h), device), "ecalj_create")
        self.set_precision(precision)

    def close(self):
        """Free every device buffer allocated through this handle, then the handle."""
        if self._h:
            for a in list(self._arrays):
                a.free()
            lib().ecalj_destroy(self._h)
            self._h = C.c_void_p()

    def array(self, shape, dtype=np.complex128):
        return DevArray(self, shape, dtype)

    def to_dev(self, a, dtype=None):
        return DevArray.from_host(self, a, dtype)

    def dev_copy(self, dst, src, nbytes=None):
        """Plain copy kernel on the handle's stream (asynchronous)."""
        n = dst.nbytes if nbytes is None else nbytes
        _chk(lib().ecalj_dev_copy(self._h, _t_ptr(dst), _t_ptr(src), n), "dev_copy")

    def memcpy_d2d(self, dst, src, nbytes=None):
        n = dst.nbytes if nbytes is None else nbytes
        assert nbytes is not None or dst.nbytes == src.nbytes
        _chk(lib().ecalj_memcpy_d2d(self._h, _t_ptr(dst), _t_ptr(src), n), "memcpy_d2d")

    def phase_similarity(self, a, ngb, nmat, d, nk, out):
        _chk(lib().ecalj_dev_phase_similarity(self._h, _t_ptr(a), ngb, nmat, _t_ptr(d), nk, _t_ptr(out)),
             "dev_phase_similarity")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ptr(self):
        return self._h

    # ---- state
    def set_precision(self, precision: str):
        """"f64" (complex128, FP64 MFMA), "f32" (complex64 on FP32 MFMA),
        "bf16x3" (complex64, hi/lo bf16 split products, FP32 accumulation) or
        "bf16f8" (bf16x3, with the Sigma_c contraction's cross terms on the
        scaled fp8 MFMA).  The epsilon inversion stays FP64 in every form."""
        if precision not in PREC:
            raise ValueError(f"precision {precision!r}: one of {sorted(PREC)}")
        self.precision = precision
        _chk(lib().ecalj_set_precision(self._h, PREC[precision]), "set_precision")

    def set_stream(self, stream_ptr: int | None):
        _chk(lib().ecalj_set_stream(self._h, C.c_void_p(stream_ptr) if stream_ptr else None), "set_stream")

    def sync(self):
        _chk(lib().ecalj_sync(self._h), "sync")

    def stream_wait(self, other: "Handle"):
        """This handle's stream waits (on the device) for the work issued so far on other's."""
        _chk(lib().ecalj_stream_wait(self._h, other._h), "stream_wait")

    def kernel_timing(self, on: bool):
        _chk(lib().ecalj_kernel_timing(self._h, int(on)), "kernel_timing")

    def kernel_time(self, kclass: str):
        ms = C.c_double()
        n = C.c_int64()
        _chk(lib().ecalj_kernel_time(self._h, KCLASS[kclass], C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value

    def graph_begin(self):
        """Record the following _dev calls into a HIP graph (after a warm-up run of the same calls)."""
        _chk(lib().ecalj_graph_begin(self._h), "graph_begin")

    def graph_end(self) -> Graph:
        g = C.c_void_p()
        _chk(lib().ecalj_graph_end(self._h, C.byref(g)), "graph_end")
        return Graph(self, g.value)

    def set_tuning(self, mode: int):
        """0: measure kernel variants on the first call of a shape; k > 0: force variant k."""
        _chk(lib().ecalj_set_tuning(self._h, int(mode)), "set_tuning")

    def tuning_report(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        _chk(lib().ecalj_tuning_report(self._h, buf, len(buf)), "tuning_report")
        return buf.value.decode()

    # ---- learned chi0 tail (no reference counterpart; see include/ecalj_gpu.h)
    # MLP batches are row-major (batch x width): a DevArray of Fortran shape (width, batch).
    def mlp_forward_dev(self, dims, params, x, batch, y):
        d = _i32(dims)
        _chk(lib().ecalj_mlp_forward_dev(self._h, len(d) - 1, _np_ptr(d), _t_ptr(params), _t_ptr(x), batch,
                                         _t_ptr(y)), "mlp_forward_dev")

    def mlp_grad_dev(self, dims, params, x, target, batch, grad, loss=None):
        d = _i32(dims)
        _chk(lib().ecalj_mlp_grad_dev(self._h, len(d) - 1, _np_ptr(d), _t_ptr(params), _t_ptr(x), _t_ptr(target),
                                      batch, _t_ptr(grad), _t_ptr(loss)), "mlp_grad_dev")

    def mlp_adam_dev(self, nparam, params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _chk(lib().ecalj_mlp_adam_dev(self._h, nparam, _t_ptr(params), _t_ptr(grad), _t_ptr(m), _t_ptr(v), step,
                                      lr, beta1, beta2, eps), "mlp_adam_dev")

    def mlp_to_bf16_dev(self, nparam, params, hi, lo=None):
        _chk(lib().ecalj_mlp_to_bf16_dev(self._h, nparam, _t_ptr(params), _t_ptr(hi), _t_ptr(lo)), "mlp_to_bf16_dev")

    def mlp_infer_bf16_dev(self, dims, params, w_hi, w_lo, x, batch, y):
        d = _i32(dims)
        _chk(lib().ecalj_mlp_infer_bf16_dev(self._h, len(d) - 1, _np_ptr(d), _t_ptr(params), _t_ptr(w_hi),
                                            _t_ptr(w_lo), _t_ptr(x), batch, _t_ptr(y)), "mlp_infer_bf16_dev")

    def tail_features_dev(self, chi, ni, nw, iref, mu, sd, feat, aux):
        _chk(lib().ecalj_tail_features_dev(self._h, _t_ptr(chi), ni, nw, iref, _t_ptr(mu), _t_ptr(sd), _t_ptr(feat),
                                           _t_ptr(aux)), "tail_features_dev")

    def tail_apply_dev(self, chi, ni, nw, iref, pred, mu_y, sd_y, aux, chi_out):
        _chk(lib().ecalj_tail_apply_dev(self._h, _t_ptr(chi), ni, nw, iref, _t_ptr(pred), _t_ptr(mu_y),
                                        _t_ptr(sd_y), _t_ptr(aux), _t_ptr(chi_out)), "tail_apply_dev")

    def set_ppb(self, nv, mdim, ppb):
        """set_ppb (S/m_zmel.f90:110-145): per-atom real blocks nv x nv x mdim."""
        nv, mdim = _i32(nv), _i32(mdim)
        flat = np.ascontiguousarray(np.asarray(ppb, np.float64).ravel(order="F"))
        assert flat.size == int((nv.astype(np.int64) ** 2 * mdim).sum())
        _chk(lib().ecalj_set_ppb(self._h, len(nv), _np_ptr(nv), _np_ptr(mdim), _np_ptr(flat)), "set_ppb")
        self.ngb = int(mdim.sum())
        self.natom = len(nv)

    def set_coulomb(self, vcoud, zcousq):
        """Readvcoud + set_m2e_prod_basis (S/m_readVcoud.f90:11-43)."""
        v = np.ascontiguousarray(np.asarray(vcoud, np.float64))
        z = _f(zcousq, np.complex128)
        _chk(lib().ecalj_set_coulomb(self._h, len(v), _np_ptr(v), _np_ptr(z)), "set_coulomb")

    def set_ppb_full(self, nlnmc, nlnmv, nblocha, ppb_full):
        """set_ppb from the reference's ppb(nlnmx, nlnmx, mdimx, natom) with core functions first."""
        c, v, m = _i32(nlnmc), _i32(nlnmv), _i32(nblocha)
        p = np.asfortranarray(ppb_full, np.float64)
        nlnmx, _, mdimx, natom = p.shape
        _chk(lib().ecalj_set_ppb_full(self._h, natom, nlnmx, mdimx, _np_ptr(c), _np_ptr(v), _np_ptr(m),
                                      _np_ptr(p)), "set_ppb_full")
        self.natom = natom
        self.ngb = int(m.sum())

    def zmel_init(self, ns1, ns2, nctot, cphim, cphiq, phase=None, iatomp=None, is_m_basis=True, zmelconjg=False):
        """get_zmel_init_gemm's MT part over middle states ns1..ns2 (core first); host arrays."""
        cq = _f(cphiq, np.complex128)
        ndima, nq = cq.shape
        cm = _f(cphim, np.complex128) if cphim is not None and np.size(cphim) else np.zeros((ndima, 0), np.complex128)
        out = np.zeros((self.ngb, ns2 - ns1 + 1, nq), np.complex128, order="F")
        ph = self._phase(phase)
        ia = _i32(iatomp) if iatomp is not None else None
        _chk(lib().ecalj_zmel_init(self._h, _np_ptr(ph), _np_ptr(ia), ndima, ns1, ns2, nctot,
                                   _np_ptr(cm) if cm.size else None, nq, _np_ptr(cq), int(is_m_basis),
                                   int(zmelconjg), 1, _np_ptr(out)), "zmel_init")
        return out

    def zhgv(self, a, b, il=1, iu=None):
        """zhgv_h (S/m_lapack.f90:64-104): eigenpairs il..iu (1-based, ascending)
        of a z = lambda b z from the upper triangles; returns (evl, z)."""
        a = _f(a, np.complex128).copy(order="F")
        b = _f(b, np.complex128)
        n = a.shape[0]
        iu = n if iu is None else iu
        evl = np.zeros(iu - il + 1)
        _chk(lib().ecalj_zhgv(self._h, _np_ptr(a), _np_ptr(b), n, _np_ptr(evl), il, iu, a.shape[0], b.shape[0]),
             "zhgv")
        return evl, a[:, :iu - il + 1].copy(order="F")

    def zhgv_dev(self, a, b, n, evl, il=1, iu=None, lda=None, ldb=None):
        iu = n if iu is None else iu
        _chk(lib().ecalj_zhgv_dev(self._h, _t_ptr(a), _t_ptr(b), n, _t_ptr(evl), il, iu, lda or n, ldb or n),
             "zhgv_dev")

    def coulomb_from_matrix(self, vcoul, oo):
        """hvccfp0's Diagonalize_Coulomb_matrix (S/main_hvccfp0.f90:322-368,
        396-401): zhgv on -vcoul (eigenvalues of vcoul in descending order),
        vcoud = -eb, zcousq = z; then Readvcoud (set_coulomb).  Returns
        (vcoud, zcousq).  The reference's KeepPositiveCou option (-eb < 0 -> 0)
        is left to the caller: set_coulomb refuses negative vcoud."""
        eb, z = self.zhgv(-np.asarray(vcoul, np.complex128), oo)
        vcoud = -eb
        self.set_coulomb(vcoud, z)
        return vcoud, z

    # ---- zmel
    def _phase(self, phase):
        if phase is None:
            phase = np.ones(self.natom, np.complex128)
        return np.ascontiguousarray(np.asarray(phase, np.complex128))

    def zmel(self, cphim, cphiq, phase=None, is_m_basis=True, zmelconjg=False):
        """get_zmel_init_gemm (S/m_zmel.f90:146-597); cphim ndima x nm (x nbatch)."""
        cm, cq = _f(cphim, np.complex128), _f(cphiq, np.complex128)
        nb = 1 if cm.ndim == 2 else cm.shape[2]
        ndima, nm, nq = cm.shape[0], cm.shape[1], cq.shape[1]
        out = np.zeros((self.ngb, nm, nq) + ((nb,) if cm.ndim == 3 else ()), np.complex128, order="F")
        ph = self._phase(phase)
        _chk(lib().ecalj_zmel(self._h, _np_ptr(ph), ndima, nm, _np_ptr(cm), nq, _np_ptr(cq), int(is_m_basis),
                              int(zmelconjg), nb, _np_ptr(out)), "zmel")
        return out

    def zmel_dev(self, cphim, cphiq, out, nm, nq, ndima, nbatch=1, phase=None, is_m_basis=True, zmelconjg=False,
                 cphim_stride=None, cphiq_stride=None):
        ph = self._phase(phase)
        sm = ndima * nm if cphim_stride is None else cphim_stride
        sq = ndima * nq if cphiq_stride is None else cphiq_stride
        _chk(lib().ecalj_zmel_strided_dev(self._h, _np_ptr(ph), ndima, nm, _t_ptr(cphim), sm, nq, _t_ptr(cphiq),
                                          sq, int(is_m_basis), int(zmelconjg), nbatch, _t_ptr(out)), "zmel_dev")
        self._keep = ph

    # ---- chi0
    def x0gemm(self, rcxq, npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, zmel, icoun_lo, icoun_hi, iwini, iwend,
               itc, itpc, jpmc, icouini, whwc):
        """x0gemm (S/x0gemm.f90:2-97); accumulates into rcxq (Fortran complex array) in place."""
        assert rcxq.dtype == np.complex128 and rcxq.flags.f_contiguous
        z = _f(zmel, np.complex128)
        nbb, _, nqtot = z.shape
        arrs = [_i32(a) for a in (iwini, iwend, itc, itpc, jpmc, icouini)]
        w = np.ascontiguousarray(np.asarray(whwc, np.float64))
        _chk(lib().ecalj_x0gemm(self._h, _np_ptr(rcxq), npr, ipr_col, npr_col, nwhis, npm, ns1, ns2, _np_ptr(z),
                                nbb, nqtot, icoun_lo, icoun_hi, *[_np_ptr(a) for a in arrs], _np_ptr(w)), "x0gemm")
        return rcxq

    def x0gemm_batched_dev(self, rcxq, npr, nwhis, nk, ns1, ns2, zmel, zmel_stride, nbb, nqtot, ix,
                           ipr_col=1, npr_col=None, npm=1):
        npr_col = npr if npr_col is None else npr_col
        ns1, ns2 = _i32(ns1), _i32(ns2)
        arrs = [_i32(a) for a in (ix.icounkmin, ix.icounkmax)]
        lists = [_i32(a) for a in (ix.iwini, ix.iwend, ix.itc, ix.itpc, ix.jpmc, ix.icouini)]
        w = np.ascontiguousarray(ix.whwc, np.float64)
        _chk(lib().ecalj_x0gemm_batched_dev(self._h, _t_ptr(rcxq), npr, ipr_col, npr_col, nwhis, npm, nk,
                                            _np_ptr(ns1), _np_ptr(ns2), _t_ptr(zmel), zmel_stride, nbb, nqtot,
                                            _np_ptr(arrs[0]), _np_ptr(arrs[1]), len(ix.itc),
                                            *[_np_ptr(a) for a in lists], len(w), _np_ptr(w)), "x0gemm_batched_dev")

    def dpsion_imag(self, rcxq, frhis, freqi, ecut=1e10, scale_rcxq=True):
        """dpsion_chiq, imaginary axis (S/dpsion5.f90:211-232); returns zxqi, scales rcxq in place."""
        assert rcxq.dtype == np.complex128 and rcxq.flags.f_contiguous
        npr, npr_col, ns = rcxq.shape
        fr = np.ascontiguousarray(frhis, np.float64)
        fi = np.ascontiguousarray(freqi, np.float64)
        zx = np.zeros((npr, npr_col, len(fi)), np.complex128, order="F")
        _chk(lib().ecalj_dpsion_imag(self._h, _np_ptr(rcxq), npr, npr_col, ns - 1, len(fi), _np_ptr(fr), _np_ptr(fi),
                                     ecut, int(scale_rcxq), _np_ptr(zx)), "dpsion_imag")
        return zx

    def dpsion_imag_dev(self, rcxq, npr, npr_col, nwhis, frhis, freqi, zxqi, ecut=1e10, scale_rcxq=False):
        fr = np.ascontiguousarray(frhis, np.float64)
        fi = np.ascontiguousarray(freqi, np.float64)
        _chk(lib().ecalj_dpsion_imag_dev(self._h, _t_ptr(rcxq), npr, npr_col, nwhis, len(fi), _np_ptr(fr),
                                         _np_ptr(fi), ecut, int(scale_rcxq), _t_ptr(zxqi)), "dpsion_imag_dev")

    def dpsion_real(self, rcxq, frhis, freqr, ecut=1e10, rcxq_scaled=False):
        """dpsion_chiq, real axis (S/dpsion5.f90:244-256); returns zxq(:,:,0:nw), rcxq unchanged."""
        r = _f(rcxq, np.complex128)
        npr, npr_col, ns = r.shape
        fr = np.ascontiguousarray(frhis, np.float64)
        fq = np.ascontiguousarray(freqr, np.float64)
        zx = np.zeros((npr, npr_col, len(fq)), np.complex128, order="F")
        _chk(lib().ecalj_dpsion_real(self._h, _np_ptr(r), npr, npr_col, ns - 1, len(fq) - 1, _np_ptr(fr), _np_ptr(fq),
                                     ecut, int(rcxq_scaled), _np_ptr(zx)), "dpsion_real")
        return zx

    def dpsion_real_dev(self, rcxq, npr, npr_col, nwhis, frhis, freqr, zxq, ecut=1e10, rcxq_scaled=False):
        fr = np.ascontiguousarray(frhis, np.float64)
        fq = np.ascontiguousarray(freqr, np.float64)
        _chk(lib().ecalj_dpsion_real_dev(self._h, _t_ptr(rcxq), npr, npr_col, nwhis, len(fq) - 1, _np_ptr(fr),
                                         _np_ptr(fq), ecut, int(rcxq_scaled), _t_ptr(zxq)), "dpsion_real_dev")

    def chi0_freq(self, z, w):
        z, w = _f(z, np.complex128), _f(w, np.float64)
        ni, npairs = z.shape
        nf = w.shape[0]
        chi = np.zeros((ni, ni, nf), np.complex128, order="F")
        _chk(lib().ecalj_chi0_freq(self._h, _np_ptr(z), _np_ptr(w), ni, npairs, nf, _np_ptr(chi)), "chi0_freq")
        return chi

    # ---- W
    def wc_build(self, chi, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True):
        """WVIllwI body for q /= 0 (S/m_llw.f90:299-373)."""
        c = _f(chi, np.complex128)
        n, _, nw = c.shape
        wc = np.zeros((n, n, nw), np.complex128, order="F")
        _chk(lib().ecalj_wc_build(self._h, _np_ptr(c), n, nw, nspin, int(is_x0_m_basis), int(is_wc_m_basis),
                                  _np_ptr(wc), n), "wc_build")
        return wc

    def wc_build_dev(self, chi, wc, ngb, nw, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True, ldw=None):
        _chk(lib().ecalj_wc_build_dev(self._h, _t_ptr(chi), ngb, nw, nspin, int(is_x0_m_basis), int(is_wc_m_basis),
                                      _t_ptr(wc), ngb if ldw is None else ldw), "wc_build_dev")

    def wc_build_real(self, chi, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True):
        """WVRllwR body for q /= 0 (S/m_llw.f90:88-180): pivoted inversion."""
        c = _f(chi, np.complex128)
        n, _, nw = c.shape
        wc = np.zeros((n, n, nw), np.complex128, order="F")
        _chk(lib().ecalj_wc_build_real(self._h, _np_ptr(c), n, nw, nspin, int(is_x0_m_basis), int(is_wc_m_basis),
                                       _np_ptr(wc), n), "wc_build_real")
        return wc

    def wc_build_real_dev(self, chi, wc, ngb, nw, nspin=1, is_x0_m_basis=True, is_wc_m_basis=True, ldw=None):
        _chk(lib().ecalj_wc_build_real_dev(self._h, _t_ptr(chi), ngb, nw, nspin, int(is_x0_m_basis),
                                           int(is_wc_m_basis), _t_ptr(wc), ngb if ldw is None else ldw),
             "wc_build_real_dev")

    def set_ppovl(self, case):
        """m_read_ppovl tables of q (synth.IPWCase fields ggg, igggi, ggg_lo, nvgcgp2, igcgp2i, gcp_lo)."""
        i32 = lambda a: np.asfortranarray(a, np.int32)
        ggg = np.ascontiguousarray(case.ggg, np.complex128)
        igg, nvg, icg = i32(case.igggi), i32(case.nvgcgp2), i32(case.igcgp2i)
        glo, gn = i32(case.ggg_lo), i32(np.array(igg.shape))
        clo, cn = i32(case.gcp_lo), i32(np.array(icg.shape))
        _chk(lib().ecalj_set_ppovl(self._h, len(ggg), _np_ptr(ggg), _np_ptr(glo), _np_ptr(gn), _np_ptr(igg),
                                   nvg.shape[1], _np_ptr(nvg), _np_ptr(clo), _np_ptr(cn), _np_ptr(icg)), "set_ppovl")

    def _ipw_args(self, case):
        i32 = lambda a: np.asfortranarray(a, np.int32)
        self._ipw_keep = [i32(case.ngveccR), i32(case.nadd), np.ascontiguousarray(case.phase, np.complex128),
                          i32(case.ngvecpB1), i32(case.ngvecpB2)]
        return self._ipw_keep

    def zmel_ipw(self, case, zmel=None, row0=0):
        """IPW rows of zmel (S/m_zmel.f90:418-538), host arrays; returns ngc x nm x ntp0 (or fills zmel)."""
        gc, nadd, ph, g1, g2 = self._ipw_args(case)
        geq = np.asfortranarray(case.geigq, np.complex128)
        dgk = np.asfortranarray(case.dgeigqk, np.complex128)
        ngc, nm, ntp0 = gc.shape[1], dgk.shape[1], geq.shape[1]
        out = np.zeros((ngc, nm, ntp0), np.complex128, order="F") if zmel is None else zmel
        _chk(lib().ecalj_zmel_ipw(self._h, ngc, _np_ptr(gc), _np_ptr(nadd), _np_ptr(ph), geq.shape[0], g1.shape[1],
                                  _np_ptr(g1), _np_ptr(geq), ntp0, g2.shape[1], _np_ptr(g2), _np_ptr(dgk), nm,
                                  _np_ptr(out), out.shape[0], row0), "zmel_ipw")
        return out

    def zmel_ipw_dev(self, case, geigq, dgeigqk, zmel, ldz, row0=0):
        gc, nadd, ph, g1, g2 = self._ipw_args(case)
        ngpmx, ntp0 = case.geigq.shape
        nm = case.dgeigqk.shape[1]
        _chk(lib().ecalj_zmel_ipw_dev(self._h, gc.shape[1], _np_ptr(gc), _np_ptr(nadd), _np_ptr(ph), ngpmx,
                                      g1.shape[1], _np_ptr(g1), _t_ptr(geigq), ntp0, g2.shape[1], _np_ptr(g2),
                                      _t_ptr(dgeigqk), nm, _t_ptr(zmel), ldz, row0), "zmel_ipw_dev")

    def zmel_assemble_dev(self, mt, nbloch, ipw, ngc, ncol, zmel, is_m_basis=False, zmelconjg=False):
        _chk(lib().ecalj_zmel_assemble_dev(self._h, _t_ptr(mt), nbloch, _t_ptr(ipw), ngc, ncol, int(is_m_basis),
                                           int(zmelconjg), _t_ptr(zmel)), "zmel_assemble_dev")

    def wv_write_dev(self, path, rec0, w, ld):
        """Records rec0.. of a __WVI/__WVR file from device W (ld x ld x nrec, complex128)."""
        nrec = w.nbytes // (16 * ld * ld)
        _chk(lib().ecalj_wv_write_dev(self._h, str(path).encode(), rec0, nrec, ld, _t_ptr(w)), "wv_write_dev")

    def wv_read_dev(self, path, rec0, w, ld):
        nrec = w.nbytes // (16 * ld * ld)
        _chk(lib().ecalj_wv_read_dev(self._h, str(path).encode(), rec0, nrec, ld, _t_ptr(w)), "wv_read_dev")

    def wv4_write_dev(self, path, rec0, w, ld):
        """Records rec0.. of a complex(4) __WVI/__WVR file (the reference's
        mixed-precision build) from device W (ld x ld x nrec, complex128)."""
        nrec = w.nbytes // (16 * ld * ld)
        _chk(lib().ecalj_wv4_write_dev(self._h, str(path).encode(), rec0, nrec, ld, _t_ptr(w)), "wv4_write_dev")

    def wv4_read_dev(self, path, rec0, w, ld):
        nrec = w.nbytes // (16 * ld * ld)
        _chk(lib().ecalj_wv4_read_dev(self._h, str(path).encode(), rec0, nrec, ld, _t_ptr(w)), "wv4_read_dev")

    def tetwt(self, mesh, ek1, ek2, ef, frhis, ebmx=1e10, npm=1, ecore=None, matrix_linear=False):
        """gettetwt (S/m_tetwt.f90:36-256): pair lists and histogram weights,
        npm = 1 or 2, with the core states of energies ecore (Ry) appended after
        the valence bands; matrix_linear: the tetrahedron_matrix_linear split
        of each bin over the corner k-points (S/tetwt5.f90:334-336, 694-759).
        mesh: synth.TetraMesh (or any object with idtetf, qbzw, ib1bz, qlat).
        With npm = 1 the lists are (nbnbx, nqbz) and nbnb (nqbz,); with npm = 2
        they keep a trailing jpm axis."""
        e1 = np.asfortranarray(ek1, np.float64)
        e2 = np.asfortranarray(ek2, np.float64)
        nband, nqbz = e1.shape
        ec = np.ascontiguousarray(np.zeros(0) if ecore is None else ecore, np.float64)
        nctot = len(ec)
        idt = np.asfortranarray(mesh.idtetf, np.int32)
        qw = np.asfortranarray(mesh.qbzw, np.float64)
        ib1 = np.ascontiguousarray(mesh.ib1bz, np.int32)
        ql = np.asfortranarray(mesh.qlat, np.float64)
        fr = np.ascontiguousarray(frhis, np.float64)
        nbx, nht = C.c_int(0), C.c_int64(0)
        args = (self._h, nband, nqbz, idt.shape[1], _np_ptr(idt), qw.shape[1], _np_ptr(qw), _np_ptr(ib1),
                _np_ptr(e1), _np_ptr(e2), ef, ebmx, _np_ptr(ql), _np_ptr(fr), len(fr) - 1, npm, nctot,
                _np_ptr(ec) if nctot else None, int(bool(matrix_linear)))
        rc = lib().ecalj_tetwt_full(*args, 0, 0, None, None, None, None, None, None, None, C.byref(nbx),
                                    C.byref(nht))
        if rc not in (0, 1):
            _chk(rc, "tetwt")
        cap = max(nbx.value, 1)
        nbnb = np.zeros((nqbz, npm), np.int32, order="F")
        arrs = [np.zeros((cap, nqbz, npm), np.int32, order="F") for _ in range(5)]
        whw = np.zeros(max(nht.value, 1))
        _chk(lib().ecalj_tetwt_full(*args, cap, len(whw), _np_ptr(nbnb), *[_np_ptr(a) for a in arrs],
                                    _np_ptr(whw), C.byref(nbx), C.byref(nht)), "tetwt")
        if npm == 1:
            nbnb, arrs = nbnb[:, 0], [a[:, :, 0] for a in arrs]
        n1b, n2b, ihw, nhw, jhw = arrs
        return dict(nbnb=nbnb, n1b=n1b, n2b=n2b, ihw=ihw, nhw=nhw, jhw=jhw, whw=whw[:nht.value],
                    nbnbx=nbx.value, nhwtot=nht.value)

    def zminv(self, a):
        """zminv (S/m_lapack.f90:21-47): returns inv(a), partial pivoting."""
        x = np.array(a, np.complex128, order="F", copy=True)
        n = x.shape[0]
        _chk(lib().ecalj_zminv(self._h, _np_ptr(x), n, n), "zminv")
        return x

    def zminv_dev(self, a, n, nbatch=1, lda=None, stride=None):
        lda = n if lda is None else lda
        stride = lda * n if stride is None else stride
        _chk(lib().ecalj_zminv_dev(self._h, _t_ptr(a), n, lda, nbatch, stride), "zminv_dev")

    def zminv_noalloc_dev(self, a, n, nbatch=1, lda=None, stride=None):
        """zminv_dev on the device's fixed scratch array (no allocation; calls
        that overlap on one device race)."""
        lda = n if lda is None else lda
        stride = lda * n if stride is None else stride
        _chk(lib().ecalj_zminv_noalloc_dev(self._h, _t_ptr(a), n, lda, nbatch, stride), "zminv_noalloc_dev")

    # ---- the Gamma point of the W build
    def wc_build_gamma(self, chi, nspin=1, is_x0_m_basis=True, real_axis=False):
        """q = 0 body (S/m_llw.f90:98-178, 299-373 with ix = 1): W in the E basis, row/column 1 zero."""
        c = _f(chi, np.complex128)
        ngb, _, nw = c.shape
        w = np.zeros((ngb, ngb, nw), np.complex128, order="F")
        _chk(lib().ecalj_wc_build_gamma(self._h, _np_ptr(c), ngb, nw, nspin, int(is_x0_m_basis), int(real_axis),
                                        _np_ptr(w), ngb), "wc_build_gamma")
        return w

    def wc_build_gamma_dev(self, chi, ngb, nw, wc, ldw=None, nspin=1, is_x0_m_basis=True, real_axis=False):
        _chk(lib().ecalj_wc_build_gamma_dev(self._h, _t_ptr(chi), ngb, nw, nspin, int(is_x0_m_basis),
                                            int(real_axis), _t_ptr(wc), ngb if ldw is None else ldw),
             "wc_build_gamma_dev")

    def eps_head(self, chi, vcou1, nspin=1, is_x0_m_basis=True, real_axis=False, wing_iw=-1):
        """Offset point q0i (S/m_llw.f90:183-256, 378-440): llw(nw) = 1/epsinv(1,1), and the wing at wing_iw."""
        c = _f(chi, np.complex128)
        ngb, _, nw = c.shape
        llw = np.zeros(nw, np.complex128)
        wing = np.zeros(ngb - 1, np.complex128)
        _chk(lib().ecalj_eps_head(self._h, _np_ptr(c), ngb, nw, nspin, int(is_x0_m_basis), float(vcou1),
                                  int(real_axis), _np_ptr(llw), wing_iw, _np_ptr(wing)), "eps_head")
        return llw, wing

    def eps_head_dev(self, chi, ngb, nw, vcou1, llw, nspin=1, is_x0_m_basis=True, real_axis=False, wing_iw=-1,
                     wing=None):
        _chk(lib().ecalj_eps_head_dev(self._h, _t_ptr(chi), ngb, nw, nspin, int(is_x0_m_basis), float(vcou1),
                                      int(real_axis), _t_ptr(llw), wing_iw,
                                      None if wing is None else _t_ptr(wing)), "eps_head_dev")

    def wv0_patch(self, w, w0, is_wc_m_basis=True):
        """modifyWV0 (S/m_w0w0i.f90:89-140): W(1,1,iw) = w0(iw), then E -> M."""
        x = np.array(w, np.complex128, order="F", copy=True)
        ngb, _, nw = x.shape
        z0 = np.ascontiguousarray(w0, np.complex128)
        _chk(lib().ecalj_wv0_patch(self._h, _np_ptr(x), ngb, ngb, nw, _np_ptr(z0), int(is_wc_m_basis)), "wv0_patch")
        return x

    def wv0_patch_dev(self, w, ngb, nw, w0, ldw=None, is_wc_m_basis=True):
        z0 = np.ascontiguousarray(w0, np.complex128)
        _chk(lib().ecalj_wv0_patch_dev(self._h, _t_ptr(w), ngb, ngb if ldw is None else ldw, nw, _np_ptr(z0),
                                       int(is_wc_m_basis)), "wv0_patch_dev")

    # ---- eigenfunctions on the full mesh
    def set_hamindex(self, case):
        """m_hamindex state from a synth.RotationCase-like object."""
        c = case
        i32 = lambda a: np.asfortranarray(a, np.int32)
        f64 = lambda a: np.asfortranarray(a, np.float64)
        self._rot_keep = [f64(c.qtt), i32(c.iqmap), i32(c.igmap), i32(c.iqimap), i32(c.ngp), i32(c.ngvecp),
                          i32(c.ngvecprev), f64(c.symops), f64(c.shtvg), i32(c.miat), f64(c.tiat), f64(c.dlmm),
                          f64(c.plat), f64(c.qlat)]
        qtt, iqm, igm, iqim, ngp, gv, gr, so, sh, mi, ti, dl, pl, ql = self._rot_keep
        _chk(lib().ecalj_set_hamindex(self._h, qtt.shape[1], _np_ptr(qtt), _np_ptr(iqm), _np_ptr(igm), _np_ptr(iqim),
                                      _np_ptr(ngp), int(c.ngpmx), int(c.imx), _np_ptr(gv), _np_ptr(gr), so.shape[2],
                                      _np_ptr(so), _np_ptr(sh), mi.shape[0], _np_ptr(mi), _np_ptr(ti), int(c.lmxax),
                                      _np_ptr(dl), _np_ptr(pl), _np_ptr(ql)), "set_hamindex")

    def set_mnla(self, ibas_tbl, l_tbl, k_tbl, offset_tbl):
        t = [np.ascontiguousarray(a, np.int32) for a in (ibas_tbl, l_tbl, k_tbl, offset_tbl)]
        _chk(lib().ecalj_set_mnla(self._h, len(t[0]), *[_np_ptr(a) for a in t]), "set_mnla")

    def readcphi_dev(self, iq, nspc, nband, cphi_irr, ldc, cphi_out):
        q = np.ascontiguousarray(iq, np.int32)
        _chk(lib().ecalj_readcphi_dev(self._h, len(q), _np_ptr(q), nspc, nband, _t_ptr(cphi_irr), ldc,
                                      _t_ptr(cphi_out)), "readcphi_dev")

    def readgeig_dev(self, iq, nspc, nband, geig_irr, ldg, geig_out):
        q = np.ascontiguousarray(iq, np.int32)
        _chk(lib().ecalj_readgeig_dev(self._h, len(q), _np_ptr(q), nspc, nband, _t_ptr(geig_irr), ldg,
                                      _t_ptr(geig_out)), "readgeig_dev")

    def rec_read_dev(self, path, recl, rec0, step, nrec, nbytes, dst, dst_stride=None):
        """Fixed-length records (__CPHI / __GEIG) to device memory."""
        _chk(lib().ecalj_rec_read_dev(self._h, str(path).encode(), int(recl), rec0, step, nrec, int(nbytes),
                                      _t_ptr(dst), int(nbytes if dst_stride is None else dst_stride)),
             "rec_read_dev")

    # ---- Sigma_c
    def sigc_weights(self, omega, ekc, esmr, wkkr, freqx, wwx, expa, ua=1.0, ns1=1, nctot=0):
        """Imaginary-axis weights (S/m_sxcf_gemm.f90:432-458)."""
        om = np.ascontiguousarray(omega, np.float64)
        ek = np.ascontiguousarray(ekc, np.float64)
        x, wt, ex = (np.ascontiguousarray(a, np.float64) for a in (freqx, wwx, expa))
        niw, nm, ntq = len(x), len(ek), len(om)
        g = np.zeros((niw + 1, nm, ntq), order="F")
        _chk(lib().ecalj_sigc_weights(self._h, niw, ns1, ns1 + nm - 1, ntq, nctot, _np_ptr(om), _np_ptr(ek), esmr,
                                      ua, wkkr, _np_ptr(x), _np_ptr(wt), _np_ptr(ex), _np_ptr(g)), "sigc_weights")
        return g

    def sigc_weights_dev(self, nbatch, niw, nm, ntq, omega, ekc, esmr, ua, wkkr, freqx, wwx, expa, out, ns1=1,
                         nctot=0):
        _chk(lib().ecalj_sigc_weights_dev(self._h, nbatch, niw, ns1, ns1 + nm - 1, ntq, nctot, _t_ptr(omega),
                                          _t_ptr(ekc), esmr, ua, _t_ptr(wkkr), _t_ptr(freqx), _t_ptr(wwx),
                                          _t_ptr(expa), _t_ptr(out)), "sigc_weights_dev")

    def sigc(self, zmel, wc, wgtim, zsec=None):
        """Sigma_c imaginary axis + final product + clip (S/m_sxcf_gemm.f90:463-496,599-608)."""
        z, w, g = _f(zmel, np.complex128), _f(wc, np.complex128), _f(wgtim, np.float64)
        ngb, nm, ntq = z.shape
        niw = w.shape[2] - 1
        zs = np.zeros((ntq, ntq), np.complex128, order="F") if zsec is None else _f(zsec, np.complex128).copy(order="F")
        _chk(lib().ecalj_sigc(self._h, _np_ptr(z), ngb, nm, ntq, niw, _np_ptr(w), w.shape[0], _np_ptr(g),
                              _np_ptr(zs), zs.shape[0]), "sigc")
        return zs

    def sigc_dev(self, nbatch, zmel, ngb, nm, ntq, niw, wc, wgtim, zsec, ldw=None, ldz=None):
        _chk(lib().ecalj_sigc_dev(self._h, nbatch, _t_ptr(zmel), ngb, nm, ntq, niw, _t_ptr(wc),
                                  ngb if ldw is None else ldw, _t_ptr(wgtim), _t_ptr(zsec),
                                  ntq if ldz is None else ldz), "sigc_dev")

    def sigc_real(self, zmel, wr, freq_r, omega, ekc, ef, esmr, wkkr, nctot=0, ns2r=None, zsec=None):
        """Real-axis pole term (S/m_sxcf_gemm.f90:501-608), one internal k, host arrays; returns zsec."""
        z, w = _f(zmel, np.complex128), _f(wr, np.complex128)
        ngb, nm, ntq = z.shape
        nw = w.shape[2] - 1
        s = np.zeros((ntq, ntq), np.complex128, order="F") if zsec is None else _f(zsec, np.complex128).copy(order="F")
        fr, om, ek = (np.ascontiguousarray(a, np.float64) for a in (freq_r, omega, ekc))
        _chk(lib().ecalj_sigc_real(self._h, _np_ptr(z), ngb, nm, ntq, _np_ptr(w), w.shape[0], nw, _np_ptr(fr),
                                   _np_ptr(om), _np_ptr(ek), ef, esmr, nctot, nm if ns2r is None else ns2r, wkkr,
                                   _np_ptr(s), ntq), "sigc_real")
        return s

    def sigc_real_dev(self, nbatch, zmel, ngb, nm, ntq, wr, wr_stride, nw, freq_r, omega, ekc, ef, esmr, wkkr, zsec,
                      nctot=0, ns2r=None, ldw=None, ldz=None):
        # ekc: nm x nbatch, column k = the middle states of internal k (Fortran order)
        fr, om, ek, wk = (np.ascontiguousarray(np.asarray(a, np.float64).ravel(order="F"))
                          for a in (freq_r, omega, ekc, wkkr))
        _chk(lib().ecalj_sigc_real_dev(self._h, nbatch, _t_ptr(zmel), ngb, nm, ntq, _t_ptr(wr),
                                       ngb if ldw is None else ldw, wr_stride, nw, _np_ptr(fr), _np_ptr(om),
                                       _np_ptr(ek), ef, esmr, nctot, nm if ns2r is None else ns2r, _np_ptr(wk),
                                       _t_ptr(zsec), ntq if ldz is None else ldz), "sigc_real_dev")

    def sigc_full_dev(self, nbatch, zmel, ngb, nm, ntq, niw, wc, wgtim, wr, wr_stride, nw, freq_r, omega, ekc, ef,
                      esmr, wkkr, zsec, nctot=0, ns2r=None, ldw=None, ldwr=None, ldz=None):
        """Imaginary axis + real-axis pole term in one czmelwc (S/m_sxcf_gemm.f90:423-608)."""
        fr, om, ek, wk = (np.ascontiguousarray(np.asarray(a, np.float64).ravel(order="F"))
                          for a in (freq_r, omega, ekc, wkkr))
        _chk(lib().ecalj_sigc_full_dev(self._h, nbatch, _t_ptr(zmel), ngb, nm, ntq, niw, _t_ptr(wc),
                                       ngb if ldw is None else ldw, _t_ptr(wgtim), _t_ptr(wr),
                                       ngb if ldwr is None else ldwr, wr_stride, nw, _np_ptr(fr), _np_ptr(om),
                                       _np_ptr(ek), ef, esmr, nctot, nm if ns2r is None else ns2r, _np_ptr(wk),
                                       _t_ptr(zsec), ntq if ldz is None else ldz), "sigc_full_dev")

    def sigc_grouped_dev(self, nm, zmel, ngb, ntq, niw, wc, wgtim, zsec, ldw=None, ldz=None):
        """Internal k-points with their own middle-state counts nm[k]; zmel and wgtim packed per k."""
        n = _i32(nm)
        _chk(lib().ecalj_sigc_grouped_dev(self._h, len(n), _np_ptr(n), _t_ptr(zmel), ngb, ntq, niw, _t_ptr(wc),
                                          ngb if ldw is None else ldw, _t_ptr(wgtim), _t_ptr(zsec),
                                          ntq if ldz is None else ldz), "sigc_grouped_dev")


# ---------------------------------------------------------------- files shared with the Fortran steps
def vcoud_write(path, q, vcoud, zcousq):
    """__Vcoud.iq (S/main_hvccfp0.f90:396-401)."""
    v = np.ascontiguousarray(vcoud, np.float64)
    z = _f(zcousq, np.complex128)
    qq = np.ascontiguousarray(q, np.float64)
    _chk(lib().ecalj_vcoud_write(str(path).encode(), len(v), _np_ptr(qq), _np_ptr(v), _np_ptr(z)), "vcoud_write")


def vcoud_read(path):
    """Returns (q, vcoud, zcousq) of a __Vcoud.iq file (S/m_readVcoud.f90:36-43)."""
    n = C.c_int(0)
    rc = lib().ecalj_vcoud_read(str(path).encode(), C.byref(n), None, None, None, 0)
    if rc not in (0, 1):
        _chk(rc, "vcoud_read")
    q = np.zeros(3)
    v = np.zeros(n.value)
    z = np.zeros((n.value, n.value), np.complex128, order="F")
    _chk(lib().ecalj_vcoud_read(str(path).encode(), C.byref(n), _np_ptr(q), _np_ptr(v), _np_ptr(z), n.value),
         "vcoud_read")
    return q, v, z


def sec2_write(path, header8, isp, q, zsec):
    """SEC2U (S/main_hsfp0.sc.f90:339-359); zsec ntq x ntq x nq."""
    h8 = np.ascontiguousarray(header8, np.int32)
    z = _f(zsec, np.complex128)
    ntq, nq = z.shape[0], z.shape[2]
    i = np.ascontiguousarray(isp, np.int32)
    qq = np.asfortranarray(q, np.float64)
    _chk(lib().ecalj_sec2_write(str(path).encode(), _np_ptr(h8), nq, ntq, _np_ptr(i), _np_ptr(qq), _np_ptr(z)),
         "sec2_write")


def sec2_read(path):
    h8 = np.zeros(8, np.int32)
    rc = lib().ecalj_sec2_read(str(path).encode(), _np_ptr(h8), 0, 0, None, None, None)
    if rc not in (0, 1):
        _chk(rc, "sec2_read")
    nq, ntq = int(h8[1]), int(h8[2])
    i = np.zeros(nq, np.int32)
    q = np.zeros((3, nq), order="F")
    z = np.zeros((ntq, ntq, nq), np.complex128, order="F")
    _chk(lib().ecalj_sec2_read(str(path).encode(), _np_ptr(h8), nq, ntq, _np_ptr(i), _np_ptr(q), _np_ptr(z)),
         "sec2_read")
    return h8, i, q, z


def wv_d_write(path, nprecx, mrecl, nblochpmx, nwp, niw, nq, nw_i):
    v = np.array([nprecx, mrecl, nblochpmx, nwp, niw, nq, nw_i], np.int64)
    _chk(lib().ecalj_wv_d_write(str(path).encode(), _np_ptr(v)), "wv_d_write")


def wv_d_read(path):
    v = np.zeros(7, np.int64)
    _chk(lib().ecalj_wv_d_read(str(path).encode(), _np_ptr(v)), "wv_d_read")
    return v


def qp_energy(elda, sex, sexcore, sec, vxc):
    """QP energies (S/main_hqpe.sc.f90:152-161), eV."""
    s = _f(sec, np.complex128)
    e = np.zeros(len(elda))
    arr = [np.ascontiguousarray(a, np.float64) for a in (elda, sex, sexcore)]
    v = np.ascontiguousarray(vxc, np.float64)
    _chk(lib().ecalj_qp_energy(len(e), *[_np_ptr(a) for a in arr], _np_ptr(s), s.shape[0], _np_ptr(v),
                               _np_ptr(e)), "qp_energy")
    return e


def sigma_herm(sex2, sexcore2, sec, vxcmat):
    arr = [_f(a, np.complex128) for a in (sex2, sexcore2, sec, vxcmat)]
    se = np.zeros_like(arr[2], order="F")
    _chk(lib().ecalj_sigma_herm(arr[2].shape[0], *[_np_ptr(a) for a in arr], _np_ptr(se)), "sigma_herm")
    return se
