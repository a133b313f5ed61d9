"""GPU checks of the mixed-precision form (KIFMM3D_MODE_MIXED).

The mixed form has no counterpart in the reference: it is bounded against the
plain FP64 form on the same tree (the FP32 products must not move the result
by more than MIXED_VS_PLAIN relative L2) and against the FP64 direct sum (the
FMM error band of np).  Blocks of 16 right-hand sides take the FP32
matrix-core near field and tiled M2L; a remainder of 4 to 15 right-hand sides
runs zero-padded on the same FP32 matrix-core kernels (near field, S2M/S2L,
L2T/M2T) with the parent-grouped M2L, and one of 1 to 3 on the FP64 vector
kernels; the vector kernels take the FP32 parent-grouped matrix-core M2L with
FP64 near fields.
"""
import numpy as np
import pytest

import kifmm3d as K
import oracle as O
from problems import COEFS, clusters, densities, relerr, sphere, uniform_cube

pytestmark = pytest.mark.gpu

MIXED_VS_PLAIN = 1e-6
FMM_BAND = {4: 1e-2, 6: 2e-4, 8: 2e-5}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def run(kt, np_, ptsmax, spos, snor, ctr, rootlvl, nrhs, seed):
    print(f"\n[case] mixed kt={kt} np={np_} ptsmax={ptsmax} n={len(spos)} nrhs={nrhs}", flush=True)
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = spos, snor, spos
    f.ctr, f.rootlvl = ctr, rootlvl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    f.setup({"-fmm3d_np": str(np_), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": "10"})
    den = densities(len(spos), f.knl.sdof(), nrhs, seed)
    plain = f.eval(den, nrhs=nrhs, mode=K.MODE_PLAIN).reshape(nrhs, -1).copy()
    mixed = f.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, -1).copy()
    assert np.all(np.isfinite(mixed))
    worst = max(relerr(mixed[r], plain[r]) for r in range(nrhs))
    rng = np.random.default_rng(seed + 5)
    idx = rng.choice(len(spos), 100, replace=False)
    tdof = f.knl.tdof()
    d = O.direct(kt, COEFS[kt], spos, snor, den[0], spos[idx])
    ferr = relerr(mixed[0].reshape(-1, tdof)[idx], d)
    print(f"  mixed-vs-plain rel L2 {worst:.3e}   mixed fmm-vs-direct {ferr:.3e}", flush=True)
    assert worst <= MIXED_VS_PLAIN
    assert ferr <= FMM_BAND[np_]
    return f


def test_mixed_laplace_16rhs_np6():
    p = uniform_cube(60000, 21)
    run(111, 6, 150, p, None, (0.5, 0.5, 0.5), 1, 16, 21)


def test_mixed_laplace_21rhs_adaptive():
    p = clusters(8000, 22)
    run(111, 4, 60, p, None, (0.5, 0.5, 0.5), 1, 21, 22)


@pytest.mark.parametrize("nrhs", [1, 3, 8, 35])
def test_mixed_laplace_rhs_remainders(nrhs):
    """Right-hand-side counts that leave no full block (1, 3, 8) or a short
    last chunk (35 = 16 + 16 + 3): padded FP32 matrix-core blocks and the FP64
    vector-kernel fallback, whose densities come from the FP64 copy."""
    p = clusters(6000, 30 + nrhs)
    run(111, 4, 50, p, None, (0.5, 0.5, 0.5), 1, nrhs, 30 + nrhs)


def test_mixed_stokes_dl_fp32_m2l():
    p, n = sphere(3000, 23)
    run(321, 4, 40, p, n, (0.0, 0.0, 0.0), 0, 16, 23)


def test_mixed_vector_kernels_one_rhs():
    """One right-hand side of the 4x3 and 3x3 translation shapes (np 6 and 8)
    through the FP32 parent-grouped M2L."""
    p, n = sphere(6000, 24)
    run(321, 6, 60, p, n, (0.0, 0.0, 0.0), 0, 1, 24)
    c = clusters(8000, 25)
    run(511, 8, 80, c, None, (0.5, 0.5, 0.5), 1, 1, 25)


@pytest.mark.parametrize("kt,nrhs", [(111, 16), (111, 18), (511, 4), (311, 3)])
def test_mixed_near_coincident_pairs_across_leaves(kt, nrhs):
    """Pairs of points 1e-10 to 2e-9 apart on either side of splitting planes
    (multiples of 1/8, faces of every box of depth <= 3): the reference drops
    such pairs (r < 1e-8, kernel3d.cpp:21, :132) whichever leaves hold them.
    The FP32 offsets of the mixed near field cannot resolve these distances
    (1/8 of the unit cube is ~1e-8 in FP32 ulps), so the device set-up flags
    the near-field entries whose closest pair is that close (measured in
    FP64): the FP32 kernels skip them and an FP64 pass adds them (18
    right-hand sides: 16 on the FP32 matrix cores, whose flagged entries the
    FP64 pass adds, and 2 on the FP64 kernel, which takes every entry itself).
    Without that, each pair would add ~1/(4 pi 1e-8) to both targets and the
    mixed form would differ from the plain form by orders of magnitude."""
    p = uniform_cube(60000, 26)
    rng = np.random.default_rng(27)
    k = 48
    planes = rng.integers(1, 8, k) / 8.0
    axis = rng.integers(0, 3, k)
    base = rng.random((k, 3)) * 0.9 + 0.05
    gap = 10.0 ** rng.uniform(-10, np.log10(2e-9), k)
    a, b = base.copy(), base.copy()
    a[np.arange(k), axis] = planes - gap / 2
    b[np.arange(k), axis] = planes + gap / 2
    pts = np.concatenate([p, a, b])
    f = run(kt, 4, 150, pts, None, (0.5, 0.5, 0.5), 1, nrhs, 26)
    # every pair straddles two leaves, each sees the other: flagged entries
    assert f.plan_stats()["near_flagged"] >= k // 2


@pytest.mark.parametrize("depth", [1, 4])
@pytest.mark.parametrize("geom,nrhs", [("cube", 16), ("clusters", 32)])
def test_wave_tiled_m2l(geom, nrhs, depth, monkeypatch):
    """The tiled FP32 M2L with one wave per (tile, frequency) (form 4,
    spectra loads 1 or 4 halo parents ahead) computes the products of the
    block-per-tile form (form 0) in another FP32 summation order: the two
    agree to FP32 rounding on every level, and both stay within the mixed
    form's bound of the plain form.  The clustered tree has tiles with absent
    halo parents and target parents without children; 32 right-hand sides
    take two 16-wide groups."""
    monkeypatch.setenv("KIFMM3D_M2L", "wave")
    monkeypatch.setenv("KIFMM3D_M2L_DEPTH", str(depth))
    p = uniform_cube(20000, 41) if geom == "cube" else clusters(15000, 42)
    print(f"\n[case] wave M2L {geom} n={len(p)} nrhs={nrhs} depth={depth}", flush=True)
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = p, None, p
    f.ctr, f.rootlvl = (0.5, 0.5, 0.5), 1
    f.knl = K.Kernel3d(111, COEFS[111])
    f.setup({"-fmm3d_np": "6", "-fmm3d_let3d_ptsmax": "60", "-fmm3d_let3d_maxlevel": "10"})
    den = densities(len(p), 1, nrhs, 43)
    plain = f.eval(den, nrhs=nrhs).reshape(nrhs, -1).copy()
    nlev = len(f.settings(K.MODE_MIXED)["m2l"])
    assert f.settings(K.MODE_MIXED)["m2l"] == [4] * nlev  # the switch makes form 4 the default
    wave = f.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, -1).copy()
    f.set_settings(K.MODE_MIXED, 3, [0] * nlev)
    tile = f.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, -1).copy()
    e_wt = max(relerr(wave[r], tile[r]) for r in range(nrhs))
    e_wp = max(relerr(wave[r], plain[r]) for r in range(nrhs))
    e_tp = max(relerr(tile[r], plain[r]) for r in range(nrhs))
    print(f"  levels {nlev}: wave-vs-tiled {e_wt:.3e}; vs plain: wave {e_wp:.3e}, tiled {e_tp:.3e}", flush=True)
    assert np.all(np.isfinite(wave))
    assert e_wt <= 5e-7
    assert e_wp <= MIXED_VS_PLAIN
    # pinned per level, mixed with the other forms
    f.set_settings(K.MODE_MIXED, 3, [4 if l % 2 else 1 for l in range(nlev)])
    mix = f.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, -1)
    assert max(relerr(mix[r], plain[r]) for r in range(nrhs)) <= MIXED_VS_PLAIN
