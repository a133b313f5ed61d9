"""CPU oracle for the KIFMM3d evaluate pass -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
(the HIP library and its host code) never loads it.

Two parts:

* ``oracle/kifmm_oracle.c`` (built into ``oracle/_build/liboracle.so``): a C
  restatement of the reference's tree, lists, kernels, operators, FFT-based
  M2L and stage order (``fmm3d/fmm3d_eval.cpp:24-243`` and what it calls).
* ``pinv_svdrep`` below: the truncated-SVD pseudoinverse of
  ``common/svdrep.cpp:31-74`` and ``common/vecmatop.cpp:124-159``, computed
  with LAPACK ``dgesvd`` through ``scipy.linalg.svd(lapack_driver="gesvd")``,
  the same LAPACK routine the reference calls (``svdrep.cpp:48``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def build() -> str:
    """Compile the C restatement with the committed Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIBPATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, _D, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_setup.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, ctypes.c_int, _D]
        L.orc_check_matrices.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_set_pinv.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_info.argtypes = [ctypes.c_void_p, _I]
        L.orc_nodes.argtypes = [ctypes.c_void_p, _I]
        L.orc_list.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _I]
        L.orc_perm.argtypes = [ctypes.c_void_p, _I, _I]
        L.orc_eval.argtypes = [ctypes.c_void_p, _D, _D]
        L.orc_buffer.restype = ctypes.c_long
        L.orc_buffer.argtypes = [ctypes.c_void_p, ctypes.c_int, _D]
        L.orc_m2l_fft.argtypes = [ctypes.c_void_p, ctypes.c_int, _I, _D, _D]
        L.orc_kernel.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D, _D, ctypes.c_int, _D, _D]
        L.orc_kernel_dofs.argtypes = [ctypes.c_int, _I, _I]
        L.orc_kernel_mm.argtypes = [ctypes.c_int]
        L.orc_direct.argtypes = [ctypes.c_int, _D, ctypes.c_int, ctypes.c_int, _D, _D, _D, ctypes.c_int, _D, _D]
        L.orc_splpos.argtypes = [ctypes.c_int, ctypes.c_double, _D]
        L.orc_regpos.argtypes = [ctypes.c_int, ctypes.c_double, _D]
        L.orc_rfft3.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
        L.orc_irfft3.argtypes = [ctypes.c_int, ctypes.c_int, _D, _D]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Threads of OracleFMM.eval (n <= 0: all host cores).  Results are the
    same for every thread count."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return lib().orc_get_threads()


def _dp(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_I) if a is not None else None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def kernel_dofs(kt: int):
    s, t = ctypes.c_int(), ctypes.c_int()
    if lib().orc_kernel_dofs(kt, ctypes.byref(s), ctypes.byref(t)) != 0:
        raise ValueError(f"unsupported kernel type {kt}")
    return s.value, t.value


def kernel_mm(kt: int) -> int:
    r = lib().orc_kernel_mm(kt)
    if r < 0:
        raise ValueError(f"unsupported kernel type {kt}")
    return r


def kernel_matrix(kt, coefs, srcpos, srcnor, trgpos):
    """Kernel3d::kernel (fmm3d/kernel3d.cpp:116): (nt*tdof) x (ns*sdof) block."""
    sdof, tdof = kernel_dofs(kt)
    sp = _f64(srcpos, (-1, 3)); tp = _f64(trgpos, (-1, 3))
    sn = _f64(srcnor, (-1, 3)) if srcnor is not None else sp
    cf = _f64(coefs if len(coefs) else [1.0])
    out = np.zeros((sp.shape[0] * sdof, tp.shape[0] * tdof))  # transposed storage = column-major
    lib().orc_kernel(kt, _dp(cf), len(cf), sp.shape[0], _dp(sp), _dp(sn), tp.shape[0], _dp(tp), _dp(out))
    return out.T.copy()


def direct(kt, coefs, srcpos, srcnor, srcden, trgpos):
    """FP64 direct sum (fmm3d/dense3d.cpp:38-58, fmm3d/fmm3d_check.cpp:48-54)."""
    sdof, tdof = kernel_dofs(kt)
    sp = _f64(srcpos, (-1, 3)); tp = _f64(trgpos, (-1, 3))
    sn = _f64(srcnor, (-1, 3)) if srcnor is not None else sp
    sd = _f64(srcden).ravel()
    cf = _f64(coefs if len(coefs) else [1.0])
    out = np.zeros(tp.shape[0] * tdof)
    lib().orc_direct(kt, _dp(cf), len(cf), sp.shape[0], _dp(sp), _dp(sn), _dp(sd), tp.shape[0], _dp(tp), _dp(out))
    return out


def splpos(np_, R):
    n = lib().orc_splpos(np_, R, None)
    out = np.zeros((n, 3))
    lib().orc_splpos(np_, R, _dp(out))
    return out


def regpos(np_, R):
    n = lib().orc_regpos(np_, R, None)
    out = np.zeros((n, 3))
    lib().orc_regpos(np_, R, _dp(out))
    return out


def rfft3(x):
    """FFTW2 rfftwnd r2c of an (n, n, n, hm) real array -> (n, n, n//2+1, hm) complex."""
    x = _f64(x)
    n, hm = x.shape[0], x.shape[3]
    out = np.zeros((n, n, n // 2 + 1, hm, 2))
    lib().orc_rfft3(n, hm, _dp(x), _dp(out))
    return out[..., 0] + 1j * out[..., 1]


def irfft3(X, n):
    """FFTW2 rfftwnd c2r (unnormalised) of an (n, n, n//2+1, hm) complex array."""
    X = np.ascontiguousarray(X, dtype=np.complex128)
    hm = X.shape[3]
    inp = np.ascontiguousarray(np.stack([X.real, X.imag], axis=-1))
    out = np.zeros((n, n, n, hm))
    lib().orc_irfft3(n, hm, _dp(inp), _dp(out))
    return out


def pinv_svdrep(K, eps):
    """Truncated pseudoinverse as in common/svdrep.cpp:31-74 + common/vecmatop.cpp:124-159.

    Singular values are kept while sigma_i >= eps * sigma_0 (the scan stops at the
    first one below).  R = V_k diag(1/sigma) U_k^T.  Returns (R, kept rank).
    """
    import scipy.linalg

    U, S, VT = scipy.linalg.svd(np.array(K, dtype=np.float64, order="F"), full_matrices=False,
                                lapack_driver="gesvd", check_finite=True)
    cutoff = eps * S[0]
    cnt = 0
    while cnt < len(S) and abs(S[cnt]) >= cutoff:
        cnt += 1
    V = VT[:cnt].T * (1.0 / S[:cnt])
    return V @ U[:, :cnt].T, cnt


_PINV_CACHE: dict = {}

NODE_FIELDS = ["par", "chd", "p0", "p1", "p2", "depth", "tag", "sndeidx", "sextbeg", "sextnum",
               "tndeidx", "textbeg", "textnum"]


class OracleFMM:
    """Mirror of the reference FMM3d object (fmm3d/fmm3d.hpp:27-115), one RHS per eval."""

    def __init__(self, kt, coefs=(), np_=6, ptsmax=150, maxlevel=10, ctr=(0.0, 0.0, 0.0), rootlvl=0):
        self.kt = kt
        self.coefs = np.array(list(coefs) if len(coefs) else [1.0, 0.25], dtype=np.float64)
        self.np = np_
        self.sdof, self.tdof = kernel_dofs(kt)
        c = _f64(ctr)
        self.h = lib().orc_create(kt, _dp(self.coefs), len(self.coefs), np_, ptsmax, maxlevel, _dp(c), rootlvl)
        if not self.h:
            raise ValueError("orc_create rejected the parameters")
        self.info = None
        self.pinv_rank = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def _info(self):
        a = np.zeros(16, dtype=np.int32)
        lib().orc_info(self.h, _ip(a))
        keys = ["nnode", "level", "sndecnt", "sextcnt", "tndecnt", "textcnt", "plnUE", "plnUC", "plnDE",
                "plnDC", "sdof", "tdof", "sm", "tm", "C", "G"]
        return dict(zip(keys, (int(v) for v in a)))

    def setup(self, srcpos, srcnor, trgpos, pinv=None):
        sp = _f64(srcpos, (-1, 3)); tp = _f64(trgpos, (-1, 3))
        sn = _f64(srcnor, (-1, 3)) if srcnor is not None else sp
        self.ns, self.nt = sp.shape[0], tp.shape[0]
        if lib().orc_setup(self.h, self.ns, _dp(sp), _dp(sn), self.nt, _dp(tp)) != 0:
            raise ValueError("orc_setup failed: a point lies outside the root box")
        self.info = self._info()
        if pinv is None:
            pinv = self.reference_pinv()
        self.set_pinv(*pinv)
        return self

    def reference_pinv(self):
        """UC2UE / DC2DE at level 0 via LAPACK dgesvd (matmgnt3d.cpp:96-105, 163-171)."""
        key = (self.kt, tuple(self.coefs), self.np)
        if key not in _PINV_CACHE:
            i = self.info
            kuc = np.zeros((i["plnUE"], i["plnUC"]))  # storage of the column-major plnUC x plnUE
            kdc = np.zeros((i["plnDE"], i["plnDC"]))
            lib().orc_check_matrices(self.h, _dp(kuc), _dp(kdc))
            eps = 0.1 ** (self.np + 1)  # matmgnt3d.cpp:38-41
            uc2ue, r1 = pinv_svdrep(kuc.T, eps)
            dc2de, r2 = pinv_svdrep(kdc.T, eps)
            _PINV_CACHE[key] = (uc2ue, dc2de, (r1, r2))
        uc2ue, dc2de, ranks = _PINV_CACHE[key]
        self.pinv_rank = ranks
        return uc2ue, dc2de

    def check_matrices(self):
        i = self.info
        kuc = np.zeros((i["plnUE"], i["plnUC"]))
        kdc = np.zeros((i["plnDE"], i["plnDC"]))
        lib().orc_check_matrices(self.h, _dp(kuc), _dp(kdc))
        return kuc.T.copy(), kdc.T.copy()

    def set_pinv(self, uc2ue, dc2de):
        """uc2ue: plnUE x plnUC, dc2de: plnDE x plnDC (ordinary 2-D arrays)."""
        a = np.asfortranarray(uc2ue, dtype=np.float64)
        b = np.asfortranarray(dc2de, dtype=np.float64)
        lib().orc_set_pinv(self.h, a.ctypes.data_as(_D), b.ctypes.data_as(_D))

    def eval(self, srcden):
        sd = _f64(srcden).ravel()
        if sd.size != self.ns * self.sdof:
            raise ValueError("srcden has the wrong size")
        out = np.zeros(self.nt * self.tdof)
        if lib().orc_eval(self.h, _dp(sd), _dp(out)) != 0:
            raise RuntimeError("orc_eval failed")
        return out

    def nodes(self):
        a = np.zeros((self.info["nnode"], 13), dtype=np.int32)
        lib().orc_nodes(self.h, _ip(a))
        return a

    def list_csr(self, which):
        n = self.info["nnode"]
        tot = lib().orc_list(self.h, which, None, None)
        ptr = np.zeros(n + 1, dtype=np.int32)
        idx = np.zeros(max(tot, 1), dtype=np.int32)
        lib().orc_list(self.h, which, _ip(ptr), _ip(idx))
        return ptr, idx[:tot]

    def perm(self):
        sp = np.zeros(max(self.info["sextcnt"], 1), dtype=np.int32)
        tp = np.zeros(max(self.info["textcnt"], 1), dtype=np.int32)
        lib().orc_perm(self.h, _ip(sp), _ip(tp))
        return sp[: self.info["sextcnt"]], tp[: self.info["textcnt"]]

    def buffer(self, which):
        n = lib().orc_buffer(self.h, which, None)
        out = np.zeros(max(n, 1))
        lib().orc_buffer(self.h, which, _dp(out))
        return out[:n]

    def m2l_fft(self, level, idx, den):
        i = np.ascontiguousarray(idx, dtype=np.int32)
        d = _f64(den)
        out = np.zeros(self.info["plnDC"])
        if lib().orc_m2l_fft(self.h, level, _ip(i), _dp(d), _dp(out)) != 0:
            raise ValueError("offset is not a V-list offset")
        return out


def relerr(a, b):
    """||a - b||_2 / ||b||_2"""
    a = np.asarray(a).ravel(); b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
