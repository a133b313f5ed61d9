"""Octree and U/V/W/X list checks (fmm3d/let3d.cpp).

Brute-force invariants pin the oracle's lists: for every target leaf, every
source leaf is accounted for exactly once by U(leaf) + W(leaf) + the V and X
lists of the leaf and its ancestors (the FMM's partition of the interaction),
and every expansion-based pair is well separated.  The product's host tree
(kifmm3d_get_tree) must then equal the oracle's bit for bit.  These run on the
CPU: the product's setup builds its host products before it touches a device.
"""
import numpy as np
import pytest

import kifmm3d as K
import oracle as O
from problems import COEFS, clusters, sphere, uniform_cube

PAR, CHD, P0, P1, P2, DEPTH, TAG = range(7)


def subtree_leaves(nodes):
    n = len(nodes)
    out = [None] * n
    for g in range(n - 1, -1, -1):
        if nodes[g, CHD] < 0:
            out[g] = [g] if nodes[g, TAG] & 1 else []
        else:
            out[g] = [x for c in range(8) for x in out[nodes[g, CHD] + c]]
    return out


def adjacent(nodes, a, b):
    """Touching boxes (let3d.cpp:339-352) computed from box extents on the finest grid."""
    md = max(nodes[a, DEPTH], nodes[b, DEPTH])
    lo_a = nodes[a, P0:P2 + 1] << (md - nodes[a, DEPTH])
    lo_b = nodes[b, P0:P2 + 1] << (md - nodes[b, DEPTH])
    hi_a = lo_a + (1 << (md - nodes[a, DEPTH]))
    hi_b = lo_b + (1 << (md - nodes[b, DEPTH]))
    overlap_or_touch = np.all(lo_a <= hi_b) and np.all(lo_b <= hi_a)
    interior_overlap = np.all(lo_a < hi_b) and np.all(lo_b < hi_a)
    return bool(overlap_or_touch and not interior_overlap)


def check_lists(nodes, lists):
    leaves_of = subtree_leaves(nodes)
    src_leaves = [g for g in range(len(nodes)) if nodes[g, CHD] < 0 and nodes[g, TAG] & 1]
    pos = {g: i for i, g in enumerate(src_leaves)}
    (up, ui), (vp, vi), (wp, wi), (xp, xi) = lists
    for t in range(len(nodes)):
        if not (nodes[t, TAG] & 2) or nodes[t, CHD] >= 0:
            continue
        cnt = np.zeros(len(src_leaves), dtype=int)
        for s in ui[up[t]:up[t + 1]]:
            assert nodes[s, CHD] < 0 and (s == t or adjacent(nodes, s, t))
            cnt[pos[s]] += 1
        for s in wi[wp[t]:wp[t + 1]]:
            assert not adjacent(nodes, s, t) and nodes[s, DEPTH] > nodes[t, DEPTH]
            for l in leaves_of[s]:
                cnt[pos[l]] += 1
        a = t
        while a >= 0:
            for s in vi[vp[a]:vp[a + 1]]:
                assert nodes[s, DEPTH] == nodes[a, DEPTH] and not adjacent(nodes, s, a)
                assert 2 <= np.abs(nodes[s, P0:P2 + 1] - nodes[a, P0:P2 + 1]).max() <= 3
                for l in leaves_of[s]:
                    cnt[pos[l]] += 1
            for s in xi[xp[a]:xp[a + 1]]:
                assert nodes[s, CHD] < 0 and nodes[s, DEPTH] < nodes[a, DEPTH] and not adjacent(nodes, s, a)
                cnt[pos[s]] += 1
            a = nodes[a, PAR]
        assert np.all(cnt == 1), f"target leaf {t}: coverage {np.bincount(cnt)}"


CASES = [
    ("uniform", lambda: (uniform_cube(3000, 31), None), (0.5, 0.5, 0.5), 1, 30),
    ("clusters", lambda: (clusters(4000, 32), None), (0.5, 0.5, 0.5), 1, 20),
    ("sphere", lambda: sphere(3000, 33), (0.0, 0.0, 0.0), 0, 25),
]


@pytest.mark.parametrize("name,gen,ctr,rl,ptsmax", CASES, ids=[c[0] for c in CASES])
def test_oracle_lists_partition_the_interactions(name, gen, ctr, rl, ptsmax):
    p, n = gen()
    o = O.OracleFMM(111, [1.0], 4, ptsmax, 10, ctr, rl)
    o.setup(p, n, p)
    nodes = o.nodes()
    lists = [o.list_csr(w) for w in range(4)]
    if name == "clusters":
        assert len(lists[2][1]) > 0 and len(lists[3][1]) > 0
    check_lists(nodes, lists)
    # split rule (let3d.cpp:86): a leaf holds <= ptsmax sources unless at depth maxlevel-1
    leaf = nodes[:, CHD] < 0
    assert np.all(nodes[leaf, 9] <= ptsmax)
    # leaf point order: ascending caller index inside every leaf (stable distribution)
    sp, _ = o.perm()
    for g in np.where(leaf & (nodes[:, TAG] & 1 > 0))[0]:
        seg = sp[nodes[g, 8]:nodes[g, 8] + nodes[g, 9]]
        assert np.all(np.diff(seg) > 0)


def test_oracle_separate_targets_and_maxlevel():
    s = clusters(3000, 34)
    t = uniform_cube(2000, 35)
    o = O.OracleFMM(111, [1.0], 4, 10, 4, (0.5, 0.5, 0.5), 1)  # maxlevel 4 caps the depth at 3
    o.setup(s, None, t)
    nodes = o.nodes()
    assert nodes[:, DEPTH].max() == 3
    check_lists(nodes, [o.list_csr(w) for w in range(4)])


def product_host(kt, p, n, t, ctr, rl, ptsmax, np_=4, maxlevel=10):
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = p, n, t
    f.ctr, f.rootlvl = ctr, rl
    f.knl = K.Kernel3d(kt, COEFS[kt])
    opts = {"-fmm3d_np": str(np_), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": str(maxlevel)}
    try:
        f.setup(opts)
    except K.KifmmError as e:
        if e.code != K.EDEVICE:  # no GPU here: the host products are still built
            raise
    return f


@pytest.mark.parametrize("name,gen,ctr,rl,ptsmax", CASES, ids=[c[0] for c in CASES])
def test_product_tree_equals_oracle(name, gen, ctr, rl, ptsmax):
    p, n = gen()
    kt = 321 if name == "sphere" else 111
    f = product_host(kt, p, n if n is not None else p, p, ctr, rl, ptsmax)
    o = O.OracleFMM(kt, COEFS[kt], 4, ptsmax, 10, ctr, rl)
    o.setup(p, n, p)
    assert np.array_equal(f.nodes(), o.nodes())
    for w in range(4):
        a, b = f.list_csr(w), o.list_csr(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert all(np.array_equal(a, b) for a, b in zip(f.perm(), o.perm()))


def test_product_tree_separate_targets():
    s = clusters(5000, 36)
    t = uniform_cube(3000, 37)
    f = product_host(111, s, s, t, (0.5, 0.5, 0.5), 1, 15, maxlevel=5)
    o = O.OracleFMM(111, [1.0], 4, 15, 5, (0.5, 0.5, 0.5), 1)
    o.setup(s, None, t)
    assert np.array_equal(f.nodes(), o.nodes())
    for w in range(4):
        assert all(np.array_equal(a, b) for a, b in zip(f.list_csr(w), o.list_csr(w)))
