"""Octree and list construction on the GPU (octant keys, radix sorts, walk to
the leaves; U/V/W/X lists by one thread per node and a segmented sort) against
the host construction, which the CPU suite pins against the oracle's
restatement of Let3d (fmm3d/let3d.cpp:56-282).  Every exported tree array --
node table, U/V/W/X lists, source and target leaf permutations -- must be
identical, including on points that lie exactly on splitting planes or on the
root-box faces, repeated points, non-dyadic box centres and the extreme
maxlevel settings; an out-of-box point must give the same error."""
import numpy as np
import pytest

import kifmm3d as K
from problems import COEFS, clusters, sphere, uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def build(monkeypatch, builder, spos, tpos, ctr, lvl, ptsmax, maxlevel, kt=111, snor=None):
    if builder == "host":
        monkeypatch.setenv("KIFMM3D_TREE", "host")
    else:
        monkeypatch.delenv("KIFMM3D_TREE", raising=False)
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = spos, snor, tpos
    f.ctr, f.rootlvl = ctr, lvl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    f.setup({"-fmm3d_np": "4", "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": str(maxlevel)})
    return f


def same_tree(monkeypatch, name, spos, tpos, ctr, lvl, ptsmax, maxlevel, kt=111, snor=None):
    print(f"\n[case] {name} ns={len(spos)} nt={len(tpos)} ptsmax={ptsmax} maxlevel={maxlevel}", flush=True)
    h = build(monkeypatch, "host", spos, tpos, ctr, lvl, ptsmax, maxlevel, kt, snor)
    d = build(monkeypatch, "device", spos, tpos, ctr, lvl, ptsmax, maxlevel, kt, snor)
    for what in range(11):
        a, b = h.tree_array(what), d.tree_array(what)
        assert a.shape == b.shape and np.array_equal(a, b), f"{name}: tree array {what} differs"
    nodes = d.nodes()
    print(f"  nodes {len(nodes)}  levels {nodes[:, 5].max() + 1}", flush=True)
    return d


def test_uniform_same_points(monkeypatch):
    p = uniform_cube(20000, 70)
    same_tree(monkeypatch, "uniform", p, p, (0.5, 0.5, 0.5), 1, 30, 10)


def test_clusters_separate_targets(monkeypatch):
    s = clusters(12000, 71)
    t = uniform_cube(9000, 72)
    same_tree(monkeypatch, "clusters/uniform", s, t, (0.5, 0.5, 0.5), 1, 20, 10)


def test_sphere_with_normals(monkeypatch):
    p, n = sphere(15000, 73)
    same_tree(monkeypatch, "sphere", p, p, (0.0, 0.0, 0.0), 0, 25, 10, kt=321, snor=n)


def test_points_on_planes_and_faces(monkeypatch):
    # dyadic grid k/32 for k = 0..32: every coordinate lies on a splitting plane
    # of some level, and k = 32 on the upper face of the root box
    g = np.arange(33) / 32.0
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    same_tree(monkeypatch, "dyadic grid", p, p[::-1].copy(), (0.5, 0.5, 0.5), 1, 5, 8)


def test_non_dyadic_centre(monkeypatch):
    rng = np.random.default_rng(74)
    ctr = np.array([0.3, -0.1, 0.7])
    p = ctr + rng.uniform(-1.0, 1.0, (15000, 3))
    # points on the first splitting planes through the centre, and on the faces
    p[:200] = ctr
    p[200:300, 0] = ctr[0]
    p[300:400, 1] = ctr[1] + 0.5
    p[400:500] = ctr + 1.0
    p[500:600] = ctr - 1.0
    same_tree(monkeypatch, "non-dyadic centre", p, p, tuple(ctr), 0, 20, 10)


def test_repeated_points_and_level_limits(monkeypatch):
    p = uniform_cube(6000, 75)
    p[:700] = p[0]  # 700 copies of one point: a leaf at depth maxlevel-1 above ptsmax
    same_tree(monkeypatch, "repeated points", p, p, (0.5, 0.5, 0.5), 1, 10, 7)
    same_tree(monkeypatch, "maxlevel 1", p, p, (0.5, 0.5, 0.5), 1, 10, 1)
    c = clusters(8000, 76, lo=1e-5, hi=1e-4)
    same_tree(monkeypatch, "maxlevel 16", c, c, (0.5, 0.5, 0.5), 1, 3, 16)


def test_out_of_box_error_matches(monkeypatch):
    p = uniform_cube(3000, 77)
    q = p.copy()
    q[1234, 2] = 1.0 + 1e-12
    q[2000, 0] = -0.5
    msgs = []
    for builder in ("host", "device"):
        with pytest.raises(K.KifmmError) as e:
            build(monkeypatch, builder, q, p, (0.5, 0.5, 0.5), 1, 30, 10)
        msgs.append((e.value.code, str(e.value)))
        with pytest.raises(K.KifmmError) as e:
            build(monkeypatch, builder, p, q, (0.5, 0.5, 0.5), 1, 30, 10)
        msgs.append((e.value.code, str(e.value)))
    assert msgs[0] == msgs[2] and msgs[1] == msgs[3], msgs
    assert "1234" in msgs[0][1] and "source" in msgs[0][1] and "target" in msgs[1][1]


def test_deep_refinement_next_to_big_leaves(monkeypatch):
    """A tight cluster refined to depth 12 beside a sparse region of large
    leaves: long W lists (descendants of a large leaf's adjacent neighbours),
    X lists of the small boxes, and coarser U entries reached from many cells
    of the window, each of which must appear once."""
    rng = np.random.default_rng(78)
    sparse = rng.random((300, 3))
    tight = 0.52 + 1e-3 * rng.standard_normal((20000, 3))
    p = np.concatenate([sparse, np.clip(tight, 0.0, 1.0)])
    d = same_tree(monkeypatch, "cluster beside large leaves", p, p, (0.5, 0.5, 0.5), 1, 8, 14)
    lens = [np.diff(d.list_csr(l)[0]) for l in range(4)]
    print("  longest lists U/V/W/X: " + " ".join(str(int(x.max())) for x in lens), flush=True)
    assert lens[2].max() > 8 and lens[3].max() > 0
