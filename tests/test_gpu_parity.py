"""GPU parity: the HIP evaluate pass against the CPU oracle on the same inputs.

The oracle (oracle/kifmm_oracle.c) restates the reference's FMM3d::eval
(fmm3d/fmm3d_eval.cpp:24-243).  Both sides use the product's pseudoinverses
(kifmm3d_get_operator -> OracleFMM.set_pinv), so the comparison isolates the
evaluate pass; the bar is 1e-12 relative L2 in FP64.  A second check against
the FP64 direct sum (fmm3d/fmm3d_check.cpp) bounds the FMM error itself.
"""

import numpy as np
import pytest

import kifmm3d as K
import oracle as O
from problems import COEFS, clusters, densities, relerr, sphere, uniform_cube

pytestmark = pytest.mark.gpu

PARITY = 1e-12                        # plain FP64 GPU vs CPU restatement
FMM_BAND = {4: 1e-2, 6: 2e-4, 8: 2e-5}  # FMM error vs direct sum, by np ("low/mid/high accuracy")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def make_fmm(kt, np_, ptsmax, maxlevel, ctr, rootlvl, spos, snor, tpos):
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = spos, snor, tpos
    f.ctr, f.rootlvl = ctr, rootlvl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    f.setup({"-fmm3d_np": str(np_), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": str(maxlevel)})
    return f


def make_oracle(f, kt, np_, ptsmax, maxlevel, ctr, rootlvl, spos, snor, tpos):
    o = O.OracleFMM(kt, COEFS[kt], np_, ptsmax, maxlevel, ctr, rootlvl)
    o.setup(spos, snor, tpos, pinv=(f.operator(0), f.operator(1)))
    return o


def run_case(name, kt, np_, ptsmax, spos, snor, tpos, ctr, rootlvl, nrhs=1, maxlevel=10, seed=0, ncheck=100):
    print(f"\n[case] {name} kt={kt} np={np_} ptsmax={ptsmax} ns={len(spos)} nt={len(tpos)} nrhs={nrhs}", flush=True)
    f = make_fmm(kt, np_, ptsmax, maxlevel, ctr, rootlvl, spos, snor, tpos)
    sdof, tdof = f.knl.sdof(), f.knl.tdof()
    den = densities(len(spos), sdof, nrhs, seed)
    val = f.eval(den, nrhs=nrhs).reshape(nrhs, -1)
    o = make_oracle(f, kt, np_, ptsmax, maxlevel, ctr, rootlvl, spos, snor, tpos)
    worst = 0.0
    for r in range(nrhs):
        ref = o.eval(den[r])
        e = relerr(val[r], ref)
        worst = max(worst, e)
    rng = np.random.default_rng(seed + 99)
    idx = rng.choice(len(tpos), min(ncheck, len(tpos)), replace=False)
    d = O.direct(kt, COEFS[kt], spos, snor, den[0], tpos[idx])
    ferr = relerr(val[0].reshape(-1, tdof)[idx], d)
    print(f"  gpu-vs-oracle rel L2 {worst:.3e}   fmm-vs-direct {ferr:.3e}   stages {f.stage_times()}", flush=True)
    assert np.all(np.isfinite(val))
    assert worst <= PARITY, f"{name}: GPU differs from the oracle by {worst:.3e}"
    assert ferr <= FMM_BAND[np_], f"{name}: FMM error {ferr:.3e} above the np={np_} band"
    return f, o, val


def test_laplace_uniform_np4():
    p = uniform_cube(6000, 1)
    run_case("laplace-uniform", 111, 4, 40, p, None, p, (0.5, 0.5, 0.5), 1)


def test_laplace_clusters_np6_adaptive():
    p = clusters(8000, 2)
    f, o, _ = run_case("laplace-clusters", 111, 6, 30, p, None, p, (0.5, 0.5, 0.5), 1)
    _, w = o.list_csr(2)
    _, x = o.list_csr(3)
    assert len(w) > 0 and len(x) > 0, "the clustered case must exercise W and X lists"


def test_stokes_dl_sphere_np4():
    p, n = sphere(6000, 3)
    run_case("stokes-dl-sphere", 321, 4, 40, p, n, p, (0.0, 0.0, 0.0), 0)


def test_stokes_dl_sphere_np6():
    p, n = sphere(4000, 4, 0.9)
    run_case("stokes-dl-sphere-np6", 321, 6, 60, p, n, p, (0.0, 0.0, 0.0), 0)


def test_navier_clusters_np4():
    p = clusters(6000, 5)
    run_case("navier-clusters", 511, 4, 40, p, None, p, (0.5, 0.5, 0.5), 1)


def test_navier_np8():
    p = clusters(3000, 6, k=3)
    run_case("navier-np8", 511, 8, 80, p, None, p, (0.5, 0.5, 0.5), 1)


def test_multi_rhs_laplace():
    p = uniform_cube(5000, 7)
    run_case("laplace-5rhs", 111, 6, 50, p, None, p, (0.5, 0.5, 0.5), 1, nrhs=5)


def test_multi_rhs_stokes():
    p, n = sphere(3000, 8)
    run_case("stokes-3rhs", 321, 4, 40, p, n, p, (0.0, 0.0, 0.0), 0, nrhs=3)


def test_separate_targets():
    s = clusters(4000, 9)
    t = uniform_cube(3000, 10)
    run_case("laplace-src-ne-trg", 111, 4, 30, s, None, t, (0.5, 0.5, 0.5), 1)


def test_root_leaf_only():
    p = uniform_cube(50, 11)
    run_case("root-leaf", 111, 4, 100, p, None, p, (0.5, 0.5, 0.5), 1)


def test_stage_buffers_match_oracle():
    p = clusters(5000, 12)
    f, o, _ = run_case("buffers", 111, 4, 30, p, None, p, (0.5, 0.5, 0.5), 1)
    for w in range(4):
        a, b = f.buffer(w), o.buffer(w)
        e = relerr(a, b)
        print(f"  buffer {w}: rel L2 {e:.3e}")
        assert e <= PARITY


def test_repeat_eval_is_deterministic():
    p = uniform_cube(4000, 13)
    f = make_fmm(111, 4, 40, 10, (0.5, 0.5, 0.5), 1, p, None, p)
    den = densities(len(p), 1, 1, 13)
    a = f.eval(den).copy()
    b = f.eval(den).copy()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_near_field_variants_21rhs(variant, monkeypatch):
    """21 right-hand sides run as blocks of 16 + 4 + 1 through every near-field
    kernel form (LDS broadcast, scalar-cache sources, matrix cores)."""
    monkeypatch.setenv("KIFMM3D_P2P", str(variant))
    p = clusters(4000, 14)
    run_case(f"laplace-21rhs-p2p{variant}", 111, 4, 40, p, None, p, (0.5, 0.5, 0.5), 1, nrhs=21)
    # leaves of up to 150 targets: chunks of 128 plus a remainder, two waves per block
    p = uniform_cube(6000, 16)
    run_case(f"laplace-16rhs-big-leaves-p2p{variant}", 111, 4, 150, p, None, p, (0.5, 0.5, 0.5), 1, nrhs=16)


@pytest.mark.parametrize("kt", [321, 511])
def test_vector_kernels_16rhs(kt):
    p = clusters(2500, 15)
    nrm = None
    if kt == 321:
        p, nrm = sphere(2500, 15)
    ctr, lvl = ((0.0, 0.0, 0.0), 0) if kt == 321 else ((0.5, 0.5, 0.5), 1)
    run_case(f"kt{kt}-16rhs", kt, 4, 40, p, nrm, p, ctr, lvl, nrhs=16)


@pytest.mark.parametrize("kt", [121, 311, 312, 322, 521])
def test_other_kernel_types(kt):
    """The remaining kernel types of the reference behind the same API
    (kernel3d.cpp:139-369, translation kernels fmm3d_setup.cpp:41-54)."""
    if kt in (121, 322, 521):
        p, n = sphere(4000, 40 + kt % 7, 0.9)
        ctr, lvl = (0.0, 0.0, 0.0), 0
    else:
        p, n = clusters(4000, 40 + kt % 7), None
        ctr, lvl = (0.5, 0.5, 0.5), 1
    run_case(f"kt{kt}", kt, 4, 40, p, n, p, ctr, lvl, nrhs=2)


@pytest.mark.parametrize("form", ["matrix", "group", "pair"])
@pytest.mark.parametrize("kt", [111, 321, 511])
def test_m2l_forms(kt, form, monkeypatch):
    """Every M2L form (parent-grouped on the matrix cores, parent-grouped on the
    vector ALUs, per V pair) against the oracle, for the three translation
    shapes 1x1, 4x3 (Stokes double layer, sdof 4) and 3x3, with 5 right-hand
    sides so that the matrix-core tiles of 16 columns mix parents and
    right-hand sides, and an adaptive tree whose boundary parents miss
    neighbour directions."""
    monkeypatch.setenv("KIFMM3D_M2L", form)
    if kt == 321:
        p, nrm = sphere(5000, 60)
        ctr, lvl = (0.0, 0.0, 0.0), 0
    else:
        p, nrm = clusters(5000, 61), None
        ctr, lvl = (0.5, 0.5, 0.5), 1
    run_case(f"kt{kt}-m2l-{form}", kt, 4, 30, p, nrm, p, ctr, lvl, nrhs=5)


@pytest.mark.parametrize("geom,nrhs", [("cube", 16), ("clusters", 32)])
def test_wave_m2l_fp64(geom, nrhs, monkeypatch):
    """The wave-per-tile M2L on the FP64 matrix cores (form 4 of the plain
    form) against the oracle at the parity bar: a uniform tree whose tiles
    have every halo parent, and a clustered one with absent halo children
    (read from the zero slot) and target parents without children; 32
    right-hand sides take two groups of 16."""
    monkeypatch.setenv("KIFMM3D_M2L", "wave")
    p = uniform_cube(12000, 64) if geom == "cube" else clusters(8000, 65)
    f, _, _ = run_case(f"wave-fp64-{geom}", 111, 4, 30, p, None, p, (0.5, 0.5, 0.5), 1, nrhs=nrhs)
    assert set(f.settings(K.MODE_PLAIN)["m2l"]) == {4}


@pytest.mark.parametrize("kt", [121, 321, 322, 521])
def test_double_layer_near_coincident_pairs(kt):
    """Targets 1e-7 to 1e-5 from a source along the sphere's tangent plane:
    r.n cancels by up to 1e7 there, so any reordering or fusing of the three
    products of r.n (kernel3d.cpp:150, :235, :257, :353) shows up at well
    above 1e-12 of the result, carried by the huge near-field values of these
    few targets."""
    p, nrm = sphere(12000, 90)
    rng = np.random.default_rng(91)
    k = 64
    src = rng.choice(len(p), k, replace=False)
    t = rng.normal(size=(k, 3))
    t -= (t * nrm[src]).sum(1)[:, None] * nrm[src]
    t /= np.linalg.norm(t, axis=1)[:, None]
    gap = 10.0 ** rng.uniform(-7, -5, k)
    q = p[src] + gap[:, None] * t
    q /= np.linalg.norm(q, axis=1)[:, None]  # back onto the sphere, still 1e-7..1e-5 away
    tp = np.concatenate([p, q])
    run_case(f"kt{kt}-near-pairs", kt, 4, 40, p, nrm, tp, (0.0, 0.0, 0.0), 0, nrhs=2)


def test_stokes_sphere_262k_all_targets():
    """The Stokes double-layer sphere workload at full size (N = 262,144 on the
    unit sphere, normals = positions, np 6, ptsmax 100, ctr 0, rootlvl 0): the
    plain FP64 form against the oracle on every target, relative L2 and
    max-abs relative to the largest value, both <= 1e-12.  The single-core
    oracle takes about 35 s here."""
    rng = np.random.default_rng(0x4B1F0002)
    n = 262144
    ct = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - ct * ct)
    p = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
    nrm = p.copy()
    f = make_fmm(321, 6, 100, 10, (0.0, 0.0, 0.0), 0, p, nrm, p)
    den = rng.uniform(-1, 1, n * 3)
    val = f.eval(den)
    o = make_oracle(f, 321, 6, 100, 10, (0.0, 0.0, 0.0), 0, p, nrm, p)
    ref = o.eval(den)
    e2 = relerr(val, ref)
    emax = float(np.max(np.abs(val - ref)) / np.max(np.abs(ref)))
    print(f"\n  stokes sphere 262k: gpu-vs-oracle rel L2 {e2:.3e}  max-abs/max {emax:.3e}", flush=True)
    assert e2 <= PARITY and emax <= PARITY
