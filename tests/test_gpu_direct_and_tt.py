"""GPU checks of the FP64 direct sum and of the reference-style `tt` driver.

kifmm3d_direct is the library's counterpart of FMM3d::check's exact sum
(fmm3d/fmm3d_check.cpp:48-54, fmm3d/dense3d.cpp:38-58): every error against
the direct sum that the benchmark and the tests report comes from it, so it is
pinned here against the oracle's direct sum (oracle/kifmm_oracle.c
orc_direct, the reference's Kernel3d::kernel applied to all sources) for the
four kernel shapes, with coincident source-target pairs (the r < 1e-8 rule)
and with 1 and 16 right-hand sides.

The `tt` test runs the built driver (kernel-independent-fast-multipole-method_amd/tt,
fmm3d/tt.cpp:40-111) on options files written in the reference's format
(fmm3d/options_0..2: `-key value %comment` lines) and checks the relative error
it prints (the reference's FMM-norm error, fmm3d_check.cpp:61-69) against the
accuracy band of np.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import kifmm3d as K
import oracle as O
from problems import COEFS, clusters, relerr, sphere, uniform_cube

pytestmark = pytest.mark.gpu

DIRECT_BAR = 1e-13
TT = os.path.join(os.path.dirname(K.__file__), "tt")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def make(kt, p, nrm, ctr, lvl):
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = p, nrm, p
    f.ctr, f.rootlvl = ctr, lvl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    f.setup({"-fmm3d_np": "4", "-fmm3d_let3d_ptsmax": "60", "-fmm3d_let3d_maxlevel": "10"})
    return f


@pytest.mark.parametrize("nrhs", [1, 16])
@pytest.mark.parametrize("kt", [111, 121, 321, 511])
def test_direct_matches_oracle(kt, nrhs):
    """GPU FP64 direct sum at sampled targets against the oracle's direct sum.
    The targets are source points, so every target has a coincident source
    that the r < 1e-8 rule must drop (kernel3d.cpp:21, :132); a few extra
    targets sit 1e-9 from a source (dropped) and 1e-6 from one (kept)."""
    print(f"\n[case] direct kt={kt} nrhs={nrhs}", flush=True)
    if kt in (121, 321):
        p, nrm = sphere(3000, 70 + kt % 11, 0.8)
        ctr, lvl = (0.0, 0.0, 0.0), 0
    else:
        p, nrm = (clusters(3000, 71) if kt == 511 else uniform_cube(3000, 72)), None
        ctr, lvl = (0.5, 0.5, 0.5), 1
    f = make(kt, p, nrm, ctr, lvl)
    sdof, tdof = f.knl.sdof(), f.knl.tdof()
    rng = np.random.default_rng(73 + kt)
    den = rng.uniform(-1, 1, (nrhs, len(p) * sdof))
    idx = np.sort(rng.choice(len(p), 120, replace=False)).astype(np.int32)
    got = f.direct(den, idx, nrhs)
    assert got.shape == (nrhs, idx.size * tdof)
    worst = 0.0
    for r in range(nrhs):
        ref = O.direct(kt, COEFS[kt], p, nrm, den[r], p[idx])
        worst = max(worst, relerr(got[r], ref))
    # targets next to sources: 1e-9 (inside the exclusion radius) and 1e-6 away
    q = np.concatenate([p[:8] + 1e-9 * np.array([1.0, 0.0, 0.0]), p[8:16] + 1e-6 * np.array([0.0, 1.0, 0.0])])
    f2 = K.FMM3d("fmm3d_")
    f2.srcpos, f2.srcnor, f2.trgpos = p, nrm, np.concatenate([p, q])
    f2.ctr, f2.rootlvl = ctr, lvl
    f2.knl = K.Kernel3d(kt, COEFS[kt])
    f2.setup({"-fmm3d_np": "4", "-fmm3d_let3d_ptsmax": "60", "-fmm3d_let3d_maxlevel": "10"})
    idq = np.arange(len(p), len(p) + len(q), dtype=np.int32)
    gq = f2.direct(den, idq, nrhs)
    for r in range(nrhs):
        ref = O.direct(kt, COEFS[kt], p, nrm, den[r], q)
        worst = max(worst, relerr(gq[r], ref))
    print(f"  GPU direct vs oracle direct: rel L2 {worst:.3e}", flush=True)
    assert np.all(np.isfinite(got)) and np.all(np.isfinite(gq))
    assert worst <= DIRECT_BAR


def test_direct_device_equals_host_pointer_path():
    p = uniform_cube(5000, 74)
    f = make(111, p, None, (0.5, 0.5, 0.5), 1)
    den = np.random.default_rng(75).uniform(-1, 1, (3, len(p)))
    idx = np.arange(0, len(p), 37, dtype=np.int32)
    a = f.direct(den, idx, 3)
    d = K.DeviceBuffer(den.nbytes)
    d.upload(den)
    b = f.direct(d, idx, 3)
    assert np.array_equal(a, b)


# options files in the reference's format (fmm3d/options_0..2)
TT_CASES = [
    (111, 4, 1e-2), (321, 6, 2e-4), (511, 8, 2e-5),  # the reference's three option sets
    (111, 6, 2e-4), (321, 4, 1e-2), (511, 6, 2e-4),
]


@pytest.mark.parametrize("kt,np_,band", TT_CASES)
def test_tt_driver_on_reference_options(kt, np_, band, tmp_path):
    if not os.path.exists(TT):
        pytest.fail(f"{TT} is not built (run __graft_entry__.build())")
    opts = tmp_path / f"options_kt{kt}_np{np_}"
    opts.write_text(
        "-numsrc 10000\t\t\t\t\t%number of source points\n"
        "-numtrg 20000\t\t\t\t\t%number of target points\n"
        f"-kt {kt}                    %kernel type\n"
        "\n"
        f"-fmm3d_np {np_}\t\t\t\t\t\t%accuracy\n"
        "-fmm3d_let3d_ptsmax 150\t\t%max number of points in a leaf\n"
        "-fmm3d_let3d_maxlevel 10\t%max number of levels\n")
    print(f"\n[case] tt kt={kt} np={np_}", flush=True)
    r = subprocess.run([TT, str(opts)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert len(re.findall(r"fmm eval used", r.stdout)) == 3
    m = re.search(r"relative error: ([0-9.eE+-]+)", r.stdout)
    assert m, "tt printed no relative error"
    err = float(m.group(1))
    assert 0.0 < err <= band, f"tt kt={kt} np={np_}: relative error {err:.3e} above {band:.0e}"


def test_tt_rejects_missing_file(tmp_path):
    if not os.path.exists(TT):
        pytest.fail(f"{TT} is not built")
    r = subprocess.run([TT, str(tmp_path / "nope")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
