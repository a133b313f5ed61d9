"""Python mirror of the reference's FMM3d / Kernel3d interface over libkifmm3d.so.

The reference drives its evaluator like this (fmm3d/tt.cpp:85-101)::

    FMM3d* fmm = new FMM3d("fmm3d_");
    fmm->srcpos() = &srcpos;  fmm->srcnor() = &srcnor;  fmm->trgpos() = &trgpos;
    fmm->ctr() = Point3(0,0,0);  fmm->rootlvl() = 0;  fmm->knl() = knl;
    fmm->setup(opts);
    fmm->eval(srcden, trgval);

and this module keeps those names and argument meanings::

    fmm = FMM3d("fmm3d_")
    fmm.srcpos = srcpos; fmm.srcnor = srcnor; fmm.trgpos = trgpos   # (N, 3) arrays
    fmm.ctr = (0, 0, 0); fmm.rootlvl = 0; fmm.knl = Kernel3d(111, [1.0, 0.25])
    fmm.setup(opts)                      # {"-fmm3d_np": "6", "-fmm3d_let3d_ptsmax": "150", ...}
    fmm.eval(srcden, trgval)             # trgval overwritten, returns 0

Positions are (N, 3) C-order arrays, i.e. the reference's 3 x N column-major
DblNumMat.  Errors raise :class:`KifmmError` (the reference asserts).  All
numerical work runs in the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libkifmm3d.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_VP = ctypes.c_void_p

# (name, restype, argtypes) of every symbol declared in include/kifmm3d.h
SYMBOLS = [
    ("kifmm3d_create", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, _D, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _D, ctypes.c_int]),
    ("kifmm3d_setup", ctypes.c_int, [_VP, _D, _D, ctypes.c_int, _D, ctypes.c_int]),
    ("kifmm3d_evaluate", ctypes.c_int, [_VP, _D, _D, ctypes.c_int, ctypes.c_int]),
    ("kifmm3d_evaluate_device", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, ctypes.c_int]),
    ("kifmm3d_evaluate_device_noalloc", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, ctypes.c_int]),
    ("kifmm3d_stage_times", ctypes.c_int, [_VP, _D, ctypes.c_int]),
    ("kifmm3d_kernel_times", ctypes.c_int, [_VP, _D, _I, ctypes.c_int]),
    ("kifmm3d_info", ctypes.c_int, [_VP, _I, ctypes.c_int]),
    ("kifmm3d_get_tree", ctypes.c_long, [_VP, ctypes.c_int, _I, ctypes.c_long]),
    ("kifmm3d_get_operator", ctypes.c_long, [_VP, ctypes.c_int, _D, ctypes.c_long]),
    ("kifmm3d_get_buffer", ctypes.c_long, [_VP, ctypes.c_int, _D, ctypes.c_long]),
    ("kifmm3d_direct", ctypes.c_int, [_VP, _D, ctypes.c_int, _I, ctypes.c_int, _D]),
    ("kifmm3d_direct_device", ctypes.c_int, [_VP, _VP, ctypes.c_int, _I, ctypes.c_int, _D]),
    ("kifmm3d_plan_stats", ctypes.c_int, [_VP, _D, ctypes.c_int]),
    ("kifmm3d_last_error", ctypes.c_char_p, [_VP]),
    ("kifmm3d_destroy", None, [_VP]),
    ("kifmm3d_device_alloc", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_long]),
    ("kifmm3d_device_free", ctypes.c_int, [_VP]),
    ("kifmm3d_memcpy_h2d", ctypes.c_int, [_VP, _VP, ctypes.c_long]),
    ("kifmm3d_memcpy_d2h", ctypes.c_int, [_VP, _VP, ctypes.c_long]),
    ("kifmm3d_device_synchronize", ctypes.c_int, []),
    ("kifmm3d_device_count", ctypes.c_int, [_I]),
    ("kifmm3d_device_copy", ctypes.c_int, [_VP, _VP, ctypes.c_long]),
    ("kifmm3d_set_operator_cache", ctypes.c_int, [ctypes.c_char_p]),
    ("kifmm3d_operator_cache_stats", ctypes.c_int, [_I, _I]),
    ("kifmm3d_release_operators", ctypes.c_int, []),
    ("kifmm3d_set_graph", ctypes.c_int, [_VP, ctypes.c_int]),
    ("kifmm3d_tune", ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, _I, ctypes.c_int, _I]),
    ("kifmm3d_settings", ctypes.c_int, [_VP, ctypes.c_int, _I, ctypes.c_int]),
    ("kifmm3d_set_settings", ctypes.c_int, [_VP, ctypes.c_int, _I, ctypes.c_int]),
    ("kifmm3d_batch_create", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.c_int]),
    ("kifmm3d_batch_evaluate_device", ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), _I, ctypes.c_int]),
    ("kifmm3d_batch_time", ctypes.c_int, [_VP, _D]),
    ("kifmm3d_batch_last_error", ctypes.c_char_p, [_VP]),
    ("kifmm3d_batch_destroy", None, [_VP]),
]

OK, EINVAL, EBOX, ESTATE, EDEVICE, ENOMEM = 0, -1, -2, -3, -4, -5
MODE_PLAIN = 0
MODE_MIXED = 1
STAGES = ["load", "up", "near", "vlist", "xlist", "down", "eval", "save", "total"]
KCLASSES = ["gather", "s2m", "gemm_up", "p2p", "fft_fwd", "m2l", "fft_inv", "s2l", "gemm_down", "s2p", "scatter"]

# kernel ids of fmm3d/kernel3d.hpp:29-53 supported by this evaluator
KNL_LAP_S_U = 111
KNL_LAP_D_U = 121
KNL_STK_S_U = 311
KNL_STK_S_P = 312
KNL_STK_D_U = 321
KNL_STK_D_P = 322
KNL_NAV_S_U = 511
KNL_NAV_D_U = 521
_DOFS = {111: (1, 1), 121: (1, 1), 311: (3, 3), 312: (3, 1), 321: (3, 3), 322: (3, 1), 511: (3, 3), 521: (3, 3),
         301: (4, 3)}

_lib = None


class KifmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libkifmm3d.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise KifmmError(ESTATE, f"{LIBPATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIBPATH)
        for name, res, args in SYMBOLS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def set_operator_cache(directory):
    """Read/write the translation operators from/to files in `directory` (None disables)."""
    lib().kifmm3d_set_operator_cache(directory.encode() if directory else None)


def operator_cache_stats():
    """(operator sets built, operator sets loaded from disk) since the process started."""
    b, l_ = ctypes.c_int(0), ctypes.c_int(0)
    lib().kifmm3d_operator_cache_stats(ctypes.byref(b), ctypes.byref(l_))
    return b.value, l_.value


def release_operators() -> int:
    """Free the in-memory operator sets no handle uses."""
    return lib().kifmm3d_release_operators()


def device_count() -> int:
    n = ctypes.c_int(0)
    if lib().kifmm3d_device_count(ctypes.byref(n)) != 0:
        return 0
    return n.value


class DeviceBuffer:
    """A block of HBM owned by the library's allocator (for device-resident inputs)."""

    def __init__(self, nbytes: int):
        self.ptr = _VP()
        self.nbytes = int(nbytes)
        rc = lib().kifmm3d_device_alloc(ctypes.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise KifmmError(rc, "device allocation failed")

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        rc = lib().kifmm3d_memcpy_h2d(self.ptr, a.ctypes.data_as(_VP), a.nbytes)
        if rc != 0:
            raise KifmmError(rc, "host to device copy failed")

    def download(self, arr: np.ndarray):
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        rc = lib().kifmm3d_memcpy_d2h(arr.ctypes.data_as(_VP), self.ptr, arr.nbytes)
        if rc != 0:
            raise KifmmError(rc, "device to host copy failed")

    def free(self):
        if self.ptr:
            lib().kifmm3d_device_free(self.ptr)
            self.ptr = _VP()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    rc = lib().kifmm3d_device_synchronize()
    if rc != 0:
        raise KifmmError(rc, "device synchronize failed")


class Kernel3d:
    """Kernel3d(kt, coefs) of fmm3d/kernel3d.hpp:56-80."""

    def __init__(self, kt: int = -1, coefs=()):
        self.kt = int(kt)
        self.coefs = [float(c) for c in coefs]

    def sdof(self) -> int:
        return _DOFS[self.kt][0]

    def tdof(self) -> int:
        return _DOFS[self.kt][1]


def options_create(path: str) -> dict:
    """optionsCreate of fmm3d/tt.cpp:23-38: first token of each line is the key, the rest the value."""
    opts = {}
    with open(path) as f:
        for line in f:
            parts = line.strip().split(None, 1)
            if not parts:
                continue
            opts[parts[0]] = parts[1] if len(parts) > 1 else ""
    return opts


def _opt_int(opts: dict, key: str) -> int:
    if key not in opts:
        raise KifmmError(EINVAL, f"missing option {key}")
    tok = opts[key].strip().split()
    try:
        return int(tok[0])
    except (IndexError, ValueError):
        raise KifmmError(EINVAL, f"option {key} is not an integer") from None


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


class FMM3d:
    """Drop-in for the reference's FMM3d (fmm3d/fmm3d.hpp:27-115) on one GPU."""

    def __init__(self, prefix: str = "fmm3d_"):
        self.prefix = prefix
        self.srcpos = None
        self.srcnor = None
        self.trgpos = None
        self.ctr = (0.0, 0.0, 0.0)
        self.rootlvl = 0
        self.knl = Kernel3d()
        self.np = 6
        self.h = None

    # -- lifecycle ---------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            msg = lib().kifmm3d_last_error(self.h).decode()
            raise KifmmError(rc, msg)

    def setup(self, opts: dict) -> int:
        """FMM3d::setup (fmm3d/fmm3d_setup.cpp:26-63) reading -<prefix>np,
        -<prefix>let3d_ptsmax and -<prefix>let3d_maxlevel."""
        np_ = _opt_int(opts, "-" + self.prefix + "np")
        ptsmax = _opt_int(opts, "-" + self.prefix + "let3d_ptsmax")
        maxlevel = _opt_int(opts, "-" + self.prefix + "let3d_maxlevel")
        if self.srcpos is None or self.trgpos is None:
            raise KifmmError(EINVAL, "srcpos and trgpos must be set before setup")
        self.np = np_
        self.destroy()
        h = _VP()
        cf = _f64(self.knl.coefs if self.knl.coefs else [1.0])
        ctr = _f64(self.ctr)
        rc = lib().kifmm3d_create(ctypes.byref(h), self.knl.kt, cf.ctypes.data_as(_D), len(cf), np_, ptsmax,
                                  maxlevel, ctr.ctypes.data_as(_D), int(self.rootlvl))
        if rc != 0:
            raise KifmmError(rc, lib().kifmm3d_last_error(None).decode())
        self.h = h
        sp = _f64(self.srcpos, (-1, 3))
        tp = _f64(self.trgpos, (-1, 3))
        sn = _f64(self.srcnor, (-1, 3)) if self.srcnor is not None else None
        self._ns, self._nt = sp.shape[0], tp.shape[0]
        rc = lib().kifmm3d_setup(self.h, sp.ctypes.data_as(_D), sn.ctypes.data_as(_D) if sn is not None else None,
                                 self._ns, tp.ctypes.data_as(_D), self._nt)
        self._check(rc)
        return 0

    def eval(self, srcden, trgval=None, nrhs: int = 1, mode: int = MODE_PLAIN):
        """FMM3d::eval (fmm3d/fmm3d_eval.cpp:24-243).  srcden holds nrhs density
        vectors of Nsrc*sdof values each; trgval (if given) is overwritten in place.
        Returns trgval."""
        if self.h is None:
            raise KifmmError(ESTATE, "eval called before setup")
        sd = _f64(srcden).ravel()
        if sd.size != nrhs * self._ns * self.knl.sdof():
            raise KifmmError(EINVAL, "srcden has the wrong size")
        n = nrhs * self._nt * self.knl.tdof()
        out = trgval if trgval is not None else np.empty(n)
        if out.size != n or out.dtype != np.float64 or not out.flags.c_contiguous:
            raise KifmmError(EINVAL, "trgval must be a contiguous float64 array of Ntrg*tdof*nrhs values")
        self._check(lib().kifmm3d_evaluate(self.h, sd.ctypes.data_as(_D), out.ctypes.data_as(_D), nrhs, mode))
        return out

    def eval_device(self, d_srcden: DeviceBuffer, d_trgval: DeviceBuffer, nrhs: int = 1, mode: int = MODE_PLAIN):
        """Evaluate with densities and values already resident in HBM."""
        self._check(lib().kifmm3d_evaluate_device(self.h, d_srcden.ptr, d_trgval.ptr, nrhs, mode))

    def eval_device_noalloc(self, d_srcden: DeviceBuffer, d_trgval: DeviceBuffer, nrhs: int = 1,
                            mode: int = MODE_PLAIN):
        """eval_device with the work buffers in the library's static per-GPU
        scratch array: no device allocation (256 MiB of work at most)."""
        self._check(lib().kifmm3d_evaluate_device_noalloc(self.h, d_srcden.ptr, d_trgval.ptr, nrhs, mode))

    def set_graph(self, on: bool = True):
        """Capture the evaluate pass into a hipGraph and replay it on repeat calls."""
        self._check(lib().kifmm3d_set_graph(self.h, 1 if on else 0))

    def tune(self, nrhs: int = 1, mode: int = MODE_PLAIN) -> dict:
        """Time the candidate near-field kernels and per-level M2L forms on this
        tree for (nrhs, mode) and keep the fastest; returns the settings chosen."""
        st = np.full(17, -1, dtype=np.int32)
        nrun = ctypes.c_int(0)
        self._check(lib().kifmm3d_tune(self.h, nrhs, mode, st.ctypes.data_as(_I), len(st), ctypes.byref(nrun)))
        out = self.settings(mode)
        out["evaluates"] = nrun.value
        return out

    def settings(self, mode: int = MODE_PLAIN) -> dict:
        """{'near': kernel, 'm2l': [form per V-list level]} in force for a mode
        (near: 0 LDS broadcast, 1 scalar-cache sources, 2 FP64 matrix cores,
        3 FP32 matrix cores, 4 FP32 LDS broadcast; m2l: 0 tiled, 1 parent-grouped matrix-core,
        2 grouped VALU, 3 per pair)."""
        st = np.full(17, -1, dtype=np.int32)
        n = lib().kifmm3d_settings(self.h, mode, st.ctypes.data_as(_I), len(st))
        if n < 0:
            self._check(n)
        return {"near": int(st[0]), "m2l": [int(v) for v in st[1:min(n, len(st))]]}

    def set_settings(self, mode: int = MODE_PLAIN, near: int = -1, m2l=()):
        """Pin the near-field kernel and per-level M2L forms for a mode (-1: default)."""
        st = np.array([near] + list(m2l), dtype=np.int32)
        self._check(lib().kifmm3d_set_settings(self.h, mode, st.ctypes.data_as(_I), len(st)))

    def destroy(self):
        if getattr(self, "h", None):
            lib().kifmm3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def direct(self, srcden, trgidx, nrhs: int = 1):
        """FP64 direct sum at the targets trgidx (caller order) on the GPU:
        returns [nrhs, len(trgidx) * tdof]."""
        idx = np.ascontiguousarray(trgidx, dtype=np.int32)
        out = np.empty((nrhs, idx.size * self.knl.tdof()))
        if isinstance(srcden, DeviceBuffer):
            rc = lib().kifmm3d_direct_device(self.h, srcden.ptr, nrhs, idx.ctypes.data_as(_I), idx.size,
                                             out.ctypes.data_as(_D))
        else:
            sd = _f64(srcden).ravel()
            if sd.size != nrhs * self._ns * self.knl.sdof():
                raise KifmmError(EINVAL, "srcden has the wrong size")
            rc = lib().kifmm3d_direct(self.h, sd.ctypes.data_as(_D), nrhs, idx.ctypes.data_as(_I), idx.size,
                                      out.ctypes.data_as(_D))
        self._check(rc)
        return out

    def check(self, srcden, trgval, numchk: int = 20, seed: int = 0, nrhs: int = 1):
        """FMM3d::check (fmm3d/fmm3d_check.cpp:25-74): compare trgval with the FP64
        direct sum at numchk sampled targets.  Sampling is seeded and without
        replacement.  Returns (||u - u_dir|| / ||u_dir||, ||u - u_dir|| / ||u||);
        the second is the reference's printed relative error."""
        rng = np.random.default_rng(seed)
        idx = np.sort(rng.choice(self._nt, min(numchk, self._nt), replace=False)).astype(np.int32)
        d = self.direct(srcden, idx, nrhs)
        tdof = self.knl.tdof()
        v = np.asarray(trgval).reshape(nrhs, self._nt, tdof)[:, idx, :].reshape(nrhs, -1)
        en = np.linalg.norm(v - d)
        return float(en / np.linalg.norm(d)), float(en / np.linalg.norm(v))

    def plan_stats(self) -> dict:
        keys = ["near_pairs", "far_pairs", "v_pairs", "s2m_pairs", "s2l_pairs", "up_flops", "down_flops",
                "fft_fwd", "fft_inv", "C", "up_flops_mixed", "near_flagged"]
        a = np.zeros(len(keys))
        self._check(lib().kifmm3d_plan_stats(self.h, a.ctypes.data_as(_D), len(keys)))
        return dict(zip(keys, a.tolist()))

    # -- introspection -----------------------------------------------------
    def stage_times(self) -> dict:
        ms = np.zeros(len(STAGES))
        self._check(lib().kifmm3d_stage_times(self.h, ms.ctypes.data_as(_D), len(STAGES)))
        return dict(zip(STAGES, ms.tolist()))

    def kernel_times(self) -> dict:
        """{class: (ms, launch groups)} of the last evaluate."""
        ms = np.zeros(len(KCLASSES))
        calls = np.zeros(len(KCLASSES), dtype=np.int32)
        self._check(lib().kifmm3d_kernel_times(self.h, ms.ctypes.data_as(_D), calls.ctypes.data_as(_I),
                                               len(KCLASSES)))
        return {k: (float(m), int(c)) for k, m, c in zip(KCLASSES, ms, calls)}

    def info(self) -> dict:
        a = np.zeros(12, dtype=np.int32)
        self._check(lib().kifmm3d_info(self.h, a.ctypes.data_as(_I), 12))
        keys = ["nnode", "nlevel", "sndecnt", "sextcnt", "tndecnt", "textcnt", "plnUE", "plnUC", "plnDE", "plnDC",
                "sdof", "tdof"]
        return dict(zip(keys, (int(v) for v in a)))

    def tree_array(self, what: int) -> np.ndarray:
        n = lib().kifmm3d_get_tree(self.h, what, None, 0)
        if n < 0:
            self._check(int(n))
        out = np.zeros(max(n, 1), dtype=np.int32)
        lib().kifmm3d_get_tree(self.h, what, out.ctypes.data_as(_I), n)
        return out[:n]

    def nodes(self) -> np.ndarray:
        return self.tree_array(0).reshape(-1, 13)

    def list_csr(self, which: int):
        """which: 0 U, 1 V, 2 W, 3 X (let3d.hpp:70-73)."""
        return self.tree_array(1 + which), self.tree_array(5 + which)

    def perm(self):
        return self.tree_array(9), self.tree_array(10)

    def operator(self, what: int) -> np.ndarray:
        """0: UC2UE (plnUE x plnUC), 1: DC2DE (plnDE x plnDC)."""
        n = lib().kifmm3d_get_operator(self.h, what, None, 0)
        if n < 0:
            self._check(int(n))
        out = np.zeros(n)
        lib().kifmm3d_get_operator(self.h, what, out.ctypes.data_as(_D), n)
        i = self.info()
        shape = (i["plnUC"], i["plnUE"]) if what == 0 else (i["plnDC"], i["plnDE"])
        return out.reshape(shape).T.copy()  # column-major storage -> 2-D array

    def buffer(self, what: int) -> np.ndarray:
        """0 upward equivalent densities, 1 upward check values, 2 downward
        equivalent densities, 3 downward check values (first RHS of the last eval)."""
        n = lib().kifmm3d_get_buffer(self.h, what, None, 0)
        if n < 0:
            self._check(int(n))
        out = np.zeros(max(n, 1))
        r = lib().kifmm3d_get_buffer(self.h, what, out.ctypes.data_as(_D), n)
        if r < 0:
            self._check(int(r))
        return out[:n]


class Batch:
    """Independent set-up FMM3d problems evaluated together in one hipGraph
    (kifmm3d_batch_*); the handles must outlive the batch."""

    def __init__(self, fmms):
        self.fmms = list(fmms)
        arr = (_VP * len(self.fmms))(*[f.h for f in self.fmms])
        self.h = _VP()
        rc = lib().kifmm3d_batch_create(ctypes.byref(self.h), arr, len(self.fmms))
        if rc != 0:
            raise KifmmError(rc, lib().kifmm3d_batch_last_error(None).decode())

    def eval_device(self, d_srcden, d_trgval, nrhs, mode: int = MODE_PLAIN):
        n = len(self.fmms)
        src = (_VP * n)(*[b.ptr for b in d_srcden])
        trg = (_VP * n)(*[b.ptr for b in d_trgval])
        nr = (ctypes.c_int * n)(*[int(r) for r in nrhs])
        rc = lib().kifmm3d_batch_evaluate_device(self.h, src, trg, nr, mode)
        if rc != 0:
            raise KifmmError(rc, lib().kifmm3d_batch_last_error(self.h).decode())

    def time(self) -> float:
        ms = ctypes.c_double(0)
        lib().kifmm3d_batch_time(self.h, ctypes.byref(ms))
        return ms.value

    def destroy(self):
        if getattr(self, "h", None):
            lib().kifmm3d_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
