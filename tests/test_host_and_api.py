"""CPU checks of the product library: it loads and exports every symbol
include/kifmm3d.h declares; the FP64 pseudoinverses of its host setup agree
with LAPACK dgesvd (the routine the reference calls, svdrep.cpp:48); the
error paths of the C ABI return status codes (the reference asserts,
common/commoninc.hpp:59-60); the options-file reader follows fmm3d/tt.cpp:23-38.
No test here launches a GPU kernel."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import kifmm3d as K
import oracle as O
from problems import COEFS, uniform_cube

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "kifmm3d.h")).read()
    return sorted(set(re.findall(r"\b(kifmm3d_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 20
    nm = subprocess.run(["nm", "-D", "--defined-only", K.LIBPATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (kifmm3d_\w+)", nm))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = K.lib()  # binds argtypes for every symbol of the Python mirror
    assert sorted(n for n, _, _ in K.SYMBOLS) == syms
    for s in syms:
        assert getattr(lib, s)


def setup_host(kt, np_, ptsmax=40, n=2000, seed=1):
    p = uniform_cube(n, seed)
    f = K.FMM3d("fmm3d_")
    f.srcpos = f.srcnor = f.trgpos = p
    f.ctr, f.rootlvl = (0.5, 0.5, 0.5), 1
    f.knl = K.Kernel3d(kt, COEFS[kt])
    try:
        f.setup({"-fmm3d_np": str(np_), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": "10"})
    except K.KifmmError as e:
        assert e.code == K.EDEVICE
    return f, p


@pytest.mark.parametrize("kt,np_", [(111, 4), (111, 6), (321, 4), (511, 4)])
def test_product_pinv_matches_lapack(kt, np_):
    f, p = setup_host(kt, np_)
    o = O.OracleFMM(kt, COEFS[kt], np_, 40, 10, (0.5, 0.5, 0.5), 1)
    o.setup(p, None, p)
    ref_uc, ref_dc = o.reference_pinv()
    kuc, kdc = o.check_matrices()
    for P, R, A in ((f.operator(0), ref_uc, kuc), (f.operator(1), ref_dc, kdc)):
        s = np.linalg.svd(A, compute_uv=False)
        r = int(np.sum(s >= 0.1 ** (np_ + 1) * s[0]))
        err = np.linalg.norm(P - R) / np.linalg.norm(R)
        assert err < 1e-14 * s[0] / s[r - 1], err
        assert np.linalg.norm(P @ A @ P - P) / np.linalg.norm(P) < 1e-8


def test_plan_stats_and_info():
    f, p = setup_host(111, 4, ptsmax=30, n=3000)
    i = f.info()
    assert i["sextcnt"] == 3000 and i["textcnt"] == 3000 and i["plnUE"] == 56 and i["plnUC"] == 152
    st = f.plan_stats()
    # near-field pairs of the plan: U list plus the direct W/X fallbacks of
    # fmm3d_eval.cpp:161 (small source leaf) and :183 (small target leaf)
    nodes = f.nodes()
    (up, ui), (wp, wi), (xp, xi) = f.list_csr(0), f.list_csr(2), f.list_csr(3)
    pairs = 0
    for t in range(len(nodes)):
        if nodes[t, 1] >= 0 or not nodes[t, 6] & 2:
            continue
        nt = int(nodes[t, 12])
        pairs += sum(nt * int(nodes[s, 9]) for s in ui[up[t]:up[t + 1]])
        pairs += sum(nt * int(nodes[s, 9]) for s in wi[wp[t]:wp[t + 1]] if nodes[s, 1] < 0 and nodes[s, 9] < i["plnUE"])
        if nt < i["plnDC"]:
            pairs += sum(nt * int(nodes[s, 9]) for s in xi[xp[t]:xp[t + 1]])
    assert len(wi) > 0 and len(xi) > 0
    assert st["near_pairs"] == pairs
    assert st["C"] == 5 * 8 * 8
    # up pass (fmm3d_eval.cpp:69-92) at depth >= 2: plain = M2M over the 8
    # children of inner nodes + UC2UE of every node; mixed = one fused
    # (UC2UE x UE2UC_c) product per inner node + UC2UE of the leaves
    ue, uc = i["plnUE"], i["plnUC"]
    src = (nodes[:, 6] & 1 != 0) & (nodes[:, 5] >= 2)
    inner = int(np.sum(src & (nodes[:, 1] >= 0)))
    leaves = int(np.sum(src & (nodes[:, 1] < 0)))
    assert inner > 0 and leaves > 0
    assert st["up_flops"] == 2.0 * (inner * 8 * ue * uc + (inner + leaves) * ue * uc)
    assert st["up_flops_mixed"] == 2.0 * (inner * 8 * ue * ue + leaves * ue * uc)


def test_error_paths():
    L = K.lib()
    h = ctypes.c_void_p()
    ctr = np.zeros(3)
    cf = np.array([1.0, 0.3])
    D = ctypes.POINTER(ctypes.c_double)
    assert L.kifmm3d_create(ctypes.byref(h), 999, cf.ctypes.data_as(D), 2, 6, 100, 10, ctr.ctypes.data_as(D), 0) == K.EINVAL
    assert b"unsupported kernel" in L.kifmm3d_last_error(None)
    assert L.kifmm3d_create(ctypes.byref(h), 111, cf.ctypes.data_as(D), 2, 5, 100, 10, ctr.ctypes.data_as(D), 0) == K.EINVAL
    assert L.kifmm3d_create(ctypes.byref(h), 111, cf.ctypes.data_as(D), 2, 6, 100, 17, ctr.ctypes.data_as(D), 0) == K.EINVAL
    assert L.kifmm3d_create(ctypes.byref(h), 511, cf.ctypes.data_as(D), 1, 6, 100, 10, ctr.ctypes.data_as(D), 0) == K.EINVAL
    assert L.kifmm3d_create(ctypes.byref(h), 111, cf.ctypes.data_as(D), 1, 6, 100, 10, ctr.ctypes.data_as(D), 0) == K.OK
    x = np.zeros(3)
    y = np.zeros(3)
    assert L.kifmm3d_evaluate(h, x.ctypes.data_as(D), y.ctypes.data_as(D), 1, 0) == K.ESTATE
    assert b"before setup" in L.kifmm3d_last_error(h)
    assert L.kifmm3d_info(h, (ctypes.c_int * 12)(), 12) == K.ESTATE
    # a point outside the root box [-1, 1]^3 (let3d.cpp:73)
    p = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0]])
    assert L.kifmm3d_setup(h, p.ctypes.data_as(D), None, 2, p.ctypes.data_as(D), 2) == K.EBOX
    assert b"outside the root box" in L.kifmm3d_last_error(h)
    L.kifmm3d_destroy(h)
    assert L.kifmm3d_create(None, 111, cf.ctypes.data_as(D), 1, 6, 100, 10, ctr.ctypes.data_as(D), 0) == K.EINVAL


def test_batch_rejects_repeated_handle():
    """A handle owns one set of work buffers, so it may appear in a batch once:
    twice, its two problems would run on concurrent graph branches over the
    same buffers."""
    L = K.lib()
    D = ctypes.POINTER(ctypes.c_double)
    ctr = np.zeros(3)
    cf = np.array([1.0])
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.kifmm3d_create(ctypes.byref(a), 111, cf.ctypes.data_as(D), 1, 4, 50, 10, ctr.ctypes.data_as(D), 0) == K.OK
    assert L.kifmm3d_create(ctypes.byref(b), 111, cf.ctypes.data_as(D), 1, 4, 50, 10, ctr.ctypes.data_as(D), 0) == K.OK
    out = ctypes.c_void_p()
    arr = (ctypes.c_void_p * 3)(a, b, a)
    assert L.kifmm3d_batch_create(ctypes.byref(out), arr, 3) == K.EINVAL
    assert b"same handle" in L.kifmm3d_batch_last_error(None)
    assert not out.value
    # distinct handles get past that check (and stop at the missing set-up)
    arr = (ctypes.c_void_p * 2)(a, b)
    assert L.kifmm3d_batch_create(ctypes.byref(out), arr, 2) == K.ESTATE
    L.kifmm3d_destroy(a)
    L.kifmm3d_destroy(b)


def test_stokes_dl_needs_normals():
    p = uniform_cube(100, 3) * 2 - 1
    f = K.FMM3d()
    f.srcpos = f.trgpos = p
    f.knl = K.Kernel3d(321, [1.0])
    with pytest.raises(K.KifmmError) as e:
        f.setup({"-fmm3d_np": "4", "-fmm3d_let3d_ptsmax": "50", "-fmm3d_let3d_maxlevel": "10"})
    assert e.value.code == K.EINVAL and "normals" in str(e.value)


def test_options_file_and_missing_keys(tmp_path):
    # same layout as the reference's fmm3d/options_0 (key, value, % comment)
    path = tmp_path / "options"
    path.write_text("-numsrc 10000\t\t\t%number of source points\n-kt 111   %laplace\n\n"
                    "-fmm3d_np 4\t\t%np=4 low accuracy\n-fmm3d_let3d_ptsmax 150\t%max points\n"
                    "-fmm3d_let3d_maxlevel 10\t%max number of levels\n")
    opts = K.options_create(str(path))
    assert opts["-numsrc"].split()[0] == "10000" and opts["-fmm3d_np"].split()[0] == "4"
    f = K.FMM3d()
    f.srcpos = f.trgpos = uniform_cube(500, 4) * 2 - 1
    f.knl = K.Kernel3d(111, [1.0])
    try:
        f.setup(opts)
    except K.KifmmError as e:
        assert e.code == K.EDEVICE
    assert f.info()["plnUE"] == 56
    del opts["-fmm3d_let3d_maxlevel"]
    with pytest.raises(K.KifmmError) as e:
        f.setup(opts)
    assert "maxlevel" in str(e.value)


def test_operator_disk_cache_roundtrip_and_corruption(tmp_path):
    """Operators written to the cache directory are read back bit for bit; a
    corrupted file is detected by its checksum and rebuilt."""
    K.set_operator_cache(str(tmp_path))
    try:
        K.release_operators()
        b0, l0 = K.operator_cache_stats()
        f1, _ = setup_host(511, 4, n=1500, seed=3)
        ref = [f1.operator(w).copy() for w in (0, 1)]
        files = sorted(tmp_path.iterdir())
        assert len(files) == 1 and files[0].name.startswith("kifmm3d_ops_v1_kt511_np4_")
        b1, l1 = K.operator_cache_stats()
        assert (b1, l1) == (b0 + 1, l0)
        del f1
        assert K.release_operators() >= 1
        f2, _ = setup_host(511, 4, n=1500, seed=4)
        assert K.operator_cache_stats() == (b1, l1 + 1)
        for w, r in zip((0, 1), ref):
            assert np.array_equal(f2.operator(w), r)
        del f2
        K.release_operators()
        data = bytearray(files[0].read_bytes())
        data[len(data) // 2] ^= 0x40
        files[0].write_bytes(bytes(data))
        f3, _ = setup_host(511, 4, n=1500, seed=5)
        assert K.operator_cache_stats() == (b1 + 1, l1 + 1)
        for w, r in zip((0, 1), ref):
            assert np.array_equal(f3.operator(w), r)
        assert files[0].read_bytes() != bytes(data)  # rewritten after the rebuild
    finally:
        K.set_operator_cache(None)


def test_settings_pin_and_validate():
    """Autotuner settings: defaults, pinning, and the kernels / forms a problem
    cannot take (host side; tuning itself times evaluates on the device)."""
    f, p = setup_host(111, 4, ptsmax=30, n=3000)
    st = f.settings(K.MODE_PLAIN)
    nlev = len(st["m2l"])
    assert nlev >= 1
    assert st["near"] == 2 and st["m2l"] == [0] * nlev
    assert f.settings(K.MODE_MIXED)["near"] == 3
    f.set_settings(K.MODE_PLAIN, 0, [1, 2] * nlev)
    assert f.settings(K.MODE_PLAIN) == {"near": 0, "m2l": ([1, 2] * nlev)[:nlev]}
    assert f.settings(K.MODE_MIXED)["near"] == 3  # per mode
    f.set_settings(K.MODE_PLAIN, -1, [-1] * nlev)
    assert f.settings(K.MODE_PLAIN) == st
    for bad in ((K.MODE_PLAIN, 3, []), (K.MODE_PLAIN, 4, []), (K.MODE_PLAIN, 0, [3]), (7, 0, [])):
        with pytest.raises(K.KifmmError) as e:
            f.set_settings(*bad)
        assert e.value.code == K.EINVAL
    with pytest.raises(K.KifmmError) as e:
        f.tune(1, K.MODE_PLAIN)
    assert e.value.code == K.ESTATE
    g, _ = setup_host(511, 4)
    assert g.settings(K.MODE_PLAIN)["near"] == 0  # the matrix-core near field is Laplace's
    with pytest.raises(K.KifmmError) as e:
        g.set_settings(K.MODE_MIXED, 3)  # FP32 matrix-core near field: Laplace only
    assert e.value.code == K.EINVAL
    with pytest.raises(K.KifmmError) as e:
        g.set_settings(K.MODE_PLAIN, 0, [0])  # tiled M2L: scalar kernels only
    assert e.value.code == K.EINVAL
    g.set_settings(K.MODE_PLAIN, 1, [1])
    assert g.settings(K.MODE_PLAIN)["near"] == 1
    # FP32 vector near field: the mixed form's default for kernels without normals
    assert g.settings(K.MODE_MIXED)["near"] == 4
    with pytest.raises(K.KifmmError) as e:
        g.set_settings(K.MODE_PLAIN, 4)
    assert e.value.code == K.EINVAL
    h, _ = setup_host(121, 4)  # with normals: FP64 near field in both forms
    assert h.settings(K.MODE_MIXED)["near"] == 0
    with pytest.raises(K.KifmmError) as e:
        h.set_settings(K.MODE_MIXED, 4)
    assert e.value.code == K.EINVAL


def test_host_code_under_sanitizers(tmp_path):
    """The host work of the library (tree, lists, plan, operators, disk cache,
    argument checks) on uniform, clustered, sphere and on-plane point sets
    with repeated points, through the C ABI, built with AddressSanitizer and
    UndefinedBehaviorSanitizer (csrc/host_check.cpp, `make sanitize`): any
    out-of-bounds access, overflow or misuse aborts it."""
    csrc = os.path.join(os.path.dirname(K.LIBPATH), "csrc")
    subprocess.run(["make", "-s", "-j", "8", "-C", csrc, "sanitize"], check=True)
    exe = os.path.join(csrc, "_build", "host_check_san")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    assert "0 failed checks" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
