"""GPU checks of the autotuner (kifmm3d_tune / kifmm3d_set_settings).

The tuner only chooses among near-field kernels and M2L forms that compute the
same sums, so a tuned or pinned evaluate must keep the plain form's parity
with the oracle (1e-12 relative L2) and the mixed form's bound against the
plain form.  Pinned settings cover every near-field kernel and a different
M2L form on neighbouring V-list levels, whatever the tuner picks on this box.
"""
import numpy as np
import pytest

import kifmm3d as K
import oracle as O
from problems import COEFS, clusters, densities, relerr, sphere, uniform_cube

pytestmark = pytest.mark.gpu

PARITY = 1e-12
MIXED_VS_PLAIN = 1e-6


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def problem(kt, n, seed):
    if kt in (121, 321, 322, 521):
        p, nrm = sphere(n, seed, 0.9)
        return p, nrm, (0.0, 0.0, 0.0), 0
    return clusters(n, seed), None, (0.5, 0.5, 0.5), 1


def make(kt, np_, ptsmax, p, nrm, ctr, lvl):
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = p, nrm, p
    f.ctr, f.rootlvl = ctr, lvl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    f.setup({"-fmm3d_np": str(np_), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": "10"})
    return f


def oracle_vals(f, kt, np_, ptsmax, p, nrm, ctr, lvl, den, rows):
    o = O.OracleFMM(kt, COEFS[kt], np_, ptsmax, 10, ctr, lvl)
    o.setup(p, nrm, p, pinv=(f.operator(0), f.operator(1)))
    return {r: o.eval(den[r]) for r in rows}


def check_parity(name, f, den, nrhs, ref, mode=K.MODE_PLAIN):
    val = f.eval(den, nrhs=nrhs, mode=mode).reshape(nrhs, -1)
    assert np.all(np.isfinite(val))
    worst = max(relerr(val[r], ref[r]) for r in ref)
    print(f"  {name}: settings {f.settings(mode)} rel L2 vs oracle {worst:.3e}", flush=True)
    assert worst <= PARITY, f"{name}: {worst:.3e}"


@pytest.mark.parametrize("kt,nrhs", [(111, 16), (321, 2), (511, 2)])
def test_tune_plain_keeps_parity(kt, nrhs):
    print(f"\n[case] tune plain kt={kt} nrhs={nrhs}", flush=True)
    p, nrm, ctr, lvl = problem(kt, 5000 if kt == 111 else 3000, 60 + kt % 7)
    f = make(kt, 4, 40, p, nrm, ctr, lvl)
    den = densities(len(p), f.knl.sdof(), nrhs, 61)
    st = f.tune(nrhs, K.MODE_PLAIN)
    print(f"  tuned: {st}", flush=True)
    assert st["evaluates"] >= 3
    assert st["near"] in ((0, 1, 2) if kt == 111 else (0, 1))
    assert len(st["m2l"]) >= 1 and all(v in ((0, 1, 2, 4) if kt == 111 else (1, 2)) for v in st["m2l"])
    assert f.settings(K.MODE_PLAIN) == {"near": st["near"], "m2l": st["m2l"]}
    ref = oracle_vals(f, kt, 4, 40, p, nrm, ctr, lvl, den, (0, nrhs - 1))
    check_parity(f"tuned kt{kt}", f, den, nrhs, ref)


@pytest.mark.parametrize("kt", [111, 321, 511])
def test_pinned_settings_every_kernel_and_mixed_levels(kt):
    """Each near-field kernel, and M2L forms that differ between neighbouring
    V-list levels, against the oracle (plain form)."""
    nrhs = 16 if kt == 111 else 2
    print(f"\n[case] pinned settings kt={kt} nrhs={nrhs}", flush=True)
    p, nrm, ctr, lvl = problem(kt, 5000 if kt == 111 else 3000, 70 + kt % 7)
    f = make(kt, 4, 40, p, nrm, ctr, lvl)
    den = densities(len(p), f.knl.sdof(), nrhs, 71)
    nlev = len(f.settings(K.MODE_PLAIN)["m2l"])
    assert nlev >= 2, "the case must have V lists on two levels at least"
    ref = oracle_vals(f, kt, 4, 40, p, nrm, ctr, lvl, den, (0, nrhs - 1))
    forms = (0, 1, 2, 4) if kt == 111 else (1, 2)
    nears = (0, 1, 2) if kt == 111 else (0, 1)
    for k, near in enumerate(nears):
        m2l = [forms[(k + l) % len(forms)] for l in range(nlev)]
        f.set_settings(K.MODE_PLAIN, near, m2l)
        assert f.settings(K.MODE_PLAIN) == {"near": near, "m2l": m2l}
        check_parity(f"kt{kt} near {near} m2l {m2l}", f, den, nrhs, ref)


def test_tune_and_pin_mixed_laplace():
    """Mixed form, 16 RHS: tuned, and pinned to the FP64 near-field kernels
    (the FP64 densities must then be gathered too) with per-level M2L forms."""
    print("\n[case] tune mixed kt=111 nrhs=16", flush=True)
    p = uniform_cube(12000, 80)
    f = make(111, 6, 60, p, None, (0.5, 0.5, 0.5), 1)
    den = densities(len(p), 1, 16, 81)
    plain = f.eval(den, nrhs=16, mode=K.MODE_PLAIN).reshape(16, -1).copy()
    st = f.tune(16, K.MODE_MIXED)
    print(f"  tuned: {st}", flush=True)
    assert st["near"] == 3
    nlev = len(st["m2l"])
    cases = [None, (2, [(1 + l) % 2 for l in range(nlev)]), (0, [l % 3 for l in range(nlev)])]
    for c in cases:
        if c is not None:
            f.set_settings(K.MODE_MIXED, c[0], c[1])
        mixed = f.eval(den, nrhs=16, mode=K.MODE_MIXED).reshape(16, -1)
        assert np.all(np.isfinite(mixed))
        worst = max(relerr(mixed[r], plain[r]) for r in range(16))
        print(f"  settings {f.settings(K.MODE_MIXED)}: mixed-vs-plain {worst:.3e}", flush=True)
        assert worst <= MIXED_VS_PLAIN
    # the plain form's settings are untouched by the mixed form's
    again = f.eval(den, nrhs=16, mode=K.MODE_PLAIN).reshape(16, -1)
    assert np.array_equal(again, plain)


@pytest.mark.parametrize("kt", [311, 312, 511])
def test_mixed_fp32_vector_near_field(kt):
    """Mixed form of kernels without normals: the FP32 vector near field (the
    default for the Stokes and Navier single layers), the FP64 ones pinned,
    and the tuned choice, against the plain form.  The Stokes pressure (312)
    keeps the FP64 near field (in FP32 this problem measured 1.6e-6 from the
    plain form) and rejects the FP32 one."""
    print(f"\n[case] mixed vector near field kt={kt}", flush=True)
    p, nrm, ctr, lvl = problem(kt, 4000, 90 + kt % 7)
    f = make(kt, 6, 60, p, nrm, ctr, lvl)
    den = densities(len(p), 3, 4, 91)
    plain = f.eval(den, nrhs=4, mode=K.MODE_PLAIN).reshape(4, -1).copy()
    fp32 = kt != 312
    assert f.settings(K.MODE_MIXED)["near"] == (4 if fp32 else 0)
    if not fp32:
        with pytest.raises(K.KifmmError) as e:
            f.set_settings(K.MODE_MIXED, 4)
        assert e.value.code == K.EINVAL
    for near in (None, 0, 1, "tune"):
        if near == "tune":
            st = f.tune(4, K.MODE_MIXED)
            assert st["near"] in ((4, 0, 1) if fp32 else (0, 1))
        elif near is not None:
            f.set_settings(K.MODE_MIXED, near)
        mixed = f.eval(den, nrhs=4, mode=K.MODE_MIXED).reshape(4, -1)
        assert np.all(np.isfinite(mixed))
        worst = max(relerr(mixed[r], plain[r]) for r in range(4))
        print(f"  settings {f.settings(K.MODE_MIXED)}: mixed-vs-plain {worst:.3e}", flush=True)
        assert worst <= MIXED_VS_PLAIN
