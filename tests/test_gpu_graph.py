"""GPU checks of graph replay (kifmm3d_set_graph) and of batched evaluation
in one graph (kifmm3d_batch_*).

Every replay is compared with an eager evaluate of the same inputs, with eager
work and device synchronisation between replays; the kernels and their order
are the same in both, so the results must agree bit for bit.
"""
import numpy as np
import pytest

import kifmm3d as K
from problems import COEFS, clusters, densities, sphere, uniform_cube

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


def buffers(f, npts, nrhs, seed):
    den = densities(npts, f.knl.sdof(), nrhs, seed)
    d_src = K.DeviceBuffer(den.nbytes)
    d_src.upload(den)
    d_trg = K.DeviceBuffer(nrhs * npts * f.knl.tdof() * 8)
    return d_src, d_trg


def download(d, nrhs, npts, tdof):
    out = np.empty((nrhs, npts * tdof))
    d.download(out)
    return out


@pytest.mark.parametrize("mode", [K.MODE_PLAIN, K.MODE_MIXED])
def test_graph_replay_matches_eager(mode):
    p = uniform_cube(30000, 31)
    nrhs = 16
    print(f"\n[case] graph replay laplace n={len(p)} nrhs={nrhs} mode={mode}", flush=True)
    f = make(111, 6, 100, p, None, (0.5, 0.5, 0.5), 1)
    d_src, d_trg = buffers(f, len(p), nrhs, 31)
    f.eval_device(d_src, d_trg, nrhs, mode)
    eager = download(d_trg, nrhs, len(p), 1)
    f.set_graph(True)
    for rep in range(3):
        f.eval_device(d_src, d_trg, nrhs, mode)
        got = download(d_trg, nrhs, len(p), 1)
        assert np.array_equal(got, eager), f"replay {rep} differs from the eager result"
        # eager work between replays: an evaluation on other buffers, then sync
        f.set_graph(False)
        e_src, e_trg = buffers(f, len(p), 1, 40 + rep)
        f.eval_device(e_src, e_trg, 1, mode)
        K.synchronize()
        f.set_graph(True)


def test_batch_matches_individual_evaluations():
    probs = [
        (111, 4, 60, clusters(12000, 32), None, (0.5, 0.5, 0.5), 1, 8),
        (321, 4, 60, *sphere(6000, 33), (0.0, 0.0, 0.0), 0, 8),
        (511, 4, 60, uniform_cube(5000, 34), None, (0.5, 0.5, 0.5), 1, 8),
        (111, 6, 80, uniform_cube(20000, 35), None, (0.5, 0.5, 0.5), 1, 16),
    ]
    fmms, srcs, trgs, nrhs, ref = [], [], [], [], []
    for i, (kt, np_, pm, p, n, ctr, lvl, nr) in enumerate(probs):
        print(f"\n[case] batch problem {i}: kt={kt} n={len(p)} nrhs={nr}", flush=True)
        f = make(kt, np_, pm, p, n, ctr, lvl)
        s, t = buffers(f, len(p), nr, 50 + i)
        f.eval_device(s, t, nr, K.MODE_PLAIN)
        ref.append(download(t, nr, len(p), f.knl.tdof()))
        fmms.append(f); srcs.append(s); trgs.append(t); nrhs.append(nr)
    b = K.Batch(fmms)
    for rep in range(3):
        for t in trgs:  # clear the outputs so a skipped problem cannot pass
            t.upload(np.zeros(t.nbytes // 8))
        b.eval_device(srcs, trgs, nrhs, K.MODE_PLAIN)
        print(f"  batch replay {rep}: {b.time():.2f} ms", flush=True)
        for i, f in enumerate(fmms):
            got = download(trgs[i], nrhs[i], len(probs[i][3]), f.knl.tdof())
            assert np.array_equal(got, ref[i]), f"replay {rep}: problem {i} differs from its own evaluate"
        # eager work between replays
        fmms[0].eval_device(srcs[0], trgs[0], nrhs[0], K.MODE_PLAIN)
        K.synchronize()
    b.destroy()


def eager_eval(f, d_src, d_trg, nrhs, npts, mode):
    f.set_graph(False)
    f.eval_device(d_src, d_trg, nrhs, mode)
    return download(d_trg, nrhs, npts, f.knl.tdof())


@pytest.mark.parametrize("mode", [K.MODE_PLAIN, K.MODE_MIXED])
def test_graphs_after_work_buffer_growth(mode):
    """A captured graph holds the device addresses of its engines' work
    buffers.  Sequence: a batch of two problems at 8 right-hand sides, an
    eager evaluate of one member at 32 (its buffers grow and move), the batch
    at 8 again; then the member's own graph at 8, a batch at 32 behind it, and
    the member's graph again.  Every launch must equal the eager result."""
    p = uniform_cube(20000, 36)
    q = clusters(15000, 37)
    print(f"\n[case] graphs after buffer growth mode={mode}", flush=True)
    f = make(111, 4, 60, p, None, (0.5, 0.5, 0.5), 1)
    g = make(511, 4, 60, q, None, (0.5, 0.5, 0.5), 1)
    s8f, t8f = buffers(f, len(p), 8, 38)
    s8g, t8g = buffers(g, len(q), 8, 39)
    s32f, t32f = buffers(f, len(p), 32, 40)
    ref8f = eager_eval(f, s8f, t8f, 8, len(p), mode)
    ref8g = eager_eval(g, s8g, t8g, 8, len(q), mode)
    b = K.Batch([f, g])
    b.eval_device([s8f, s8g], [t8f, t8g], [8, 8], mode)
    assert np.array_equal(download(t8f, 8, len(p), 1), ref8f)
    assert np.array_equal(download(t8g, 8, len(q), 3), ref8g)
    ref32f = eager_eval(f, s32f, t32f, 32, len(p), mode)  # f's work buffers grow here
    K.synchronize()
    for t in (t8f, t8g):
        t.upload(np.zeros(t.nbytes // 8))
    b.eval_device([s8f, s8g], [t8f, t8g], [8, 8], mode)
    assert np.array_equal(download(t8f, 8, len(p), 1), ref8f), "batch replayed over freed buffers"
    assert np.array_equal(download(t8g, 8, len(q), 3), ref8g)
    # the member's own graph, then a batch that grows its buffers behind it
    h = make(111, 4, 60, p, None, (0.5, 0.5, 0.5), 1)
    r8 = eager_eval(h, s8f, t8f, 8, len(p), mode)
    h.set_graph(True)
    h.eval_device(s8f, t8f, 8, mode)
    assert np.array_equal(download(t8f, 8, len(p), 1), r8)
    b2 = K.Batch([h, g])
    b2.eval_device([s32f, s8g], [t32f, t8g], [32, 8], mode)
    assert np.array_equal(download(t32f, 32, len(p), 1), ref32f)
    t8f.upload(np.zeros(t8f.nbytes // 8))
    h.eval_device(s8f, t8f, 8, mode)
    assert np.array_equal(download(t8f, 8, len(p), 1), r8), "engine graph replayed over freed buffers"
    b.destroy()
    b2.destroy()


def test_graph_after_setup_on_the_same_handle():
    """kifmm3d_setup on a handle that holds a captured graph replaces every
    device array: the next evaluate must capture again (new tree, new points)."""
    import ctypes
    p = uniform_cube(12000, 41)
    p2 = clusters(12000, 42)
    print("\n[case] graph after re-setup", flush=True)
    f = make(111, 4, 60, p, None, (0.5, 0.5, 0.5), 1)
    s, t = buffers(f, len(p), 16, 43)
    f.set_graph(True)
    f.eval_device(s, t, 16, K.MODE_PLAIN)
    D = ctypes.POINTER(ctypes.c_double)
    a = np.ascontiguousarray(p2)
    rc = K.lib().kifmm3d_setup(f.h, a.ctypes.data_as(D), None, len(p2), a.ctypes.data_as(D), len(p2))
    assert rc == 0
    t.upload(np.zeros(t.nbytes // 8))
    f.eval_device(s, t, 16, K.MODE_PLAIN)
    got = download(t, 16, len(p2), 1)
    ref = make(111, 4, 60, p2, None, (0.5, 0.5, 0.5), 1)
    s2, t2 = buffers(ref, len(p2), 16, 43)
    ref.eval_device(s2, t2, 16, K.MODE_PLAIN)
    assert np.array_equal(got, download(t2, 16, len(p2), 1))
