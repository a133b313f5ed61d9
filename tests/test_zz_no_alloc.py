"""GPU checks of the no-allocation evaluate (kifmm3d_evaluate_device_noalloc).

Its work buffers live in the library's static per-GPU scratch array instead of
the handle's own device memory; kernels, launch shapes and right-hand-side
chunks are those of kifmm3d_evaluate_device, so both entry points must give
the same values bit for bit.  Calls follow each other with smaller inputs, so
each one runs on scratch memory the previous call left dirty.  All calls go
through the C ABI.
"""
import numpy as np
import pytest

import kifmm3d as K
from problems import COEFS, clusters, densities, relerr, sphere, uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def make(kt, np_, ptsmax, p, n, ctr, lvl):
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = p, n, p
    f.ctr, f.rootlvl = ctr, lvl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    f.setup({"-fmm3d_np": str(np_), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": "10"})
    return f


def run_both(f, npts, nrhs, mode, seed):
    """(existing entry point's values, no-allocation entry point's values)"""
    den = densities(npts, f.knl.sdof(), nrhs, seed)
    d_src = K.DeviceBuffer(den.nbytes)
    d_src.upload(den)
    nt = nrhs * npts * f.knl.tdof()
    d_ref, d_got = K.DeviceBuffer(nt * 8), K.DeviceBuffer(nt * 8)
    d_got.upload(np.full(nt, np.nan))  # every value must be written
    f.eval_device(d_src, d_ref, nrhs, mode)
    rc = K.lib().kifmm3d_evaluate_device_noalloc(f.h, d_src.ptr, d_got.ptr, nrhs, mode)
    assert rc == K.OK, K.lib().kifmm3d_last_error(f.h)
    ref, got = np.empty(nt), np.empty(nt)
    d_ref.download(ref)
    d_got.download(got)
    return ref, got


def test_noalloc_matches_evaluate_device():
    # call 1: Laplace, 6,000 points (leaves at depth 2), 16 right-hand sides, both forms
    p = uniform_cube(6000, 71)
    print(f"\n[case] noalloc laplace n={len(p)} nrhs=16 np=6", flush=True)
    f = make(111, 6, 100, p, None, (0.5, 0.5, 0.5), 1)
    for mode in (K.MODE_PLAIN, K.MODE_MIXED):
        ref, got = run_both(f, len(p), 16, mode, 71)
        assert np.array_equal(got, ref), f"mode {mode}: the no-allocation evaluate differs"
    # after a no-allocation call the per-node buffers are not the handle's
    buf = np.empty(8)
    assert K.lib().kifmm3d_get_buffer(f.h, 0, buf.ctypes.data_as(K._D), buf.size) == K.ESTATE
    # call 2, smaller: Stokes double layer on a sphere, 4,000 points, 3 RHS
    q, nq = sphere(4000, 72)
    print(f"[case] noalloc stokes-dl n={len(q)} nrhs=3 np=4", flush=True)
    g = make(321, 4, 60, q, nq, (0.0, 0.0, 0.0), 0)
    ref, got = run_both(g, len(q), 3, K.MODE_PLAIN, 72)
    assert np.array_equal(got, ref)
    # call 3, smaller still: Navier clusters (W and X lists), 2,500 points, 1 RHS, mixed
    c = clusters(2500, 73)
    print(f"[case] noalloc navier n={len(c)} nrhs=1 np=4 mixed", flush=True)
    h = make(511, 4, 40, c, None, (0.5, 0.5, 0.5), 1)
    ref, got = run_both(h, len(c), 1, K.MODE_MIXED, 73)
    assert np.array_equal(got, ref)
    # the first handle again, 5 RHS, after the array held the other problems
    print("[case] noalloc laplace again nrhs=5", flush=True)
    ref, got = run_both(f, len(p), 5, K.MODE_PLAIN, 74)
    assert np.array_equal(got, ref)


def test_noalloc_smaller_chunks():
    """20,000 points with leaves at depth 3: the spectra of 16 right-hand
    sides (2 x 512 boxes x 1008 frequencies x 16 bytes each for the
    densities and the values) exceed the array, so the right-hand sides run
    in smaller chunks than kifmm3d_evaluate_device's 16: the same sums up to
    rounding (the plain form's parity bar, the mixed form's bound)."""
    p = uniform_cube(20000, 76)
    print(f"\n[case] noalloc laplace n={len(p)} nrhs=16 np=6 (smaller chunks)", flush=True)
    f = make(111, 6, 100, p, None, (0.5, 0.5, 0.5), 1)
    ref, got = run_both(f, len(p), 16, K.MODE_PLAIN, 76)
    e = relerr(got, ref)
    refm, gotm = run_both(f, len(p), 16, K.MODE_MIXED, 76)
    em = max(relerr(a, b) for a, b in zip(gotm.reshape(16, -1), ref.reshape(16, -1)))
    print(f"  plain: noalloc vs evaluate_device {e:.3e}; mixed noalloc vs plain {em:.3e}", flush=True)
    assert e <= 1e-12
    assert em <= 1e-6 and relerr(gotm, refm) <= 1e-6


def test_noalloc_rejects_work_beyond_the_scratch_array():
    # 1M uniform points at np 6: the spectra of one right-hand side alone
    # (32768 source boxes x 1008 frequencies x 16 bytes) exceed 256 MiB
    p = uniform_cube(1 << 20, 75)
    print(f"\n[case] noalloc oversize laplace n={len(p)}", flush=True)
    f = make(111, 6, 150, p, None, (0.5, 0.5, 0.5), 1)
    d_src, d_trg = K.DeviceBuffer(len(p) * 8), K.DeviceBuffer(len(p) * 8)
    rc = K.lib().kifmm3d_evaluate_device_noalloc(f.h, d_src.ptr, d_trg.ptr, 1, K.MODE_PLAIN)
    assert rc == K.ENOMEM
    assert b"scratch" in K.lib().kifmm3d_last_error(f.h)
