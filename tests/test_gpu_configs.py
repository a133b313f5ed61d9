"""The benchmark workloads at their full sizes on the GPU.

Each workload is generated exactly as bench.py generates it (same generators,
seeds 0x4B1F0000 + workload number, same kernels, np, leaf sizes and boxes):

* Laplace single layer, 65,536 uniform points, np 4, ptsmax 100: the plain
  form against the oracle on every target (<= 1e-12), the mixed form within
  1e-6 of it, both inside the np 4 band against the direct sum.
* Navier single layer, 1,048,576 points in 32 Gaussian clusters, np 8,
  ptsmax 150: the plain form against the oracle on every target (<= 2e-12,
  the oracle's own summation-order floor on this workload; max-abs / max
  <= 1e-12), the mixed form within 1e-6 of the plain form.
* Laplace single layer, 4,194,304 uniform points x 16 right-hand sides, np 6,
  ptsmax 256 (the leaf size the benchmark's search picks): mixed within 1e-6
  of plain on every value, both against the GPU direct sum on 1000 targets x
  16 right-hand sides, which sits at the np 6 truncation error of this
  workload (~1.3e-6); np 8 on the same points falls below 1e-6.
* The 64-problem batch (8 right-hand sides each) in the mixed form in one
  graph: every problem's batch output bitwise equal to its own eager
  evaluate, each problem inside its np band against the direct sum, and the
  worst Stokes problem of each np traced to the algorithm (the oracle shows
  the same error).

The error variant of the reference's check (FMM-norm denominator,
fmm3d/fmm3d_check.cpp:61-69) is printed beside the direct-sum-norm one.
"""
import time

import numpy as np
import pytest

import bench
import kifmm3d as K
import oracle as O
from problems import relerr

pytestmark = pytest.mark.gpu

PARITY = 1e-12
# Navier clusters, 1,048,576 points, np 8, 9 levels: the oracle itself moves by
# 1.1e-12 to 1.7e-12 when any one of its summation orders (kernel sums, dense
# mat-vecs, V-list accumulation, DFT sums) is reversed (measured on this
# workload; tests/test_oracle_numerics.py::test_rounding_floor_navier_clusters_1m
# measures the kernel-sum one), so the reference's result is defined only to
# that level here and the bar is set at the floor
PARITY_DEEP = 2e-12
MIXED_VS_PLAIN = 1e-6
FMM_BAND = {4: 1e-2, 6: 2e-4, 8: 2e-5}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if K.device_count() < 1:
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")


def setup(c, pos, nor, ptsmax, np_=None):
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = pos, nor, pos
    f.ctr, f.rootlvl = c["ctr"], c["rootlvl"]
    f.knl = K.Kernel3d(c["kt"], c["coefs"])
    f.setup({"-fmm3d_np": str(np_ or c["np"]), "-fmm3d_let3d_ptsmax": str(ptsmax), "-fmm3d_let3d_maxlevel": "10"})
    return f


def errors(val, d):
    """(direct-sum-norm error, FMM-norm error of fmm3d_check.cpp:61-69)"""
    e = np.linalg.norm(val - d)
    return float(e / np.linalg.norm(d)), float(e / np.linalg.norm(val))


def check_vs_direct(f, den, nrhs, n, tdof, nchk, seed):
    rng = np.random.default_rng(seed + 2)
    idx = np.sort(rng.choice(n, nchk, replace=False)).astype(np.int32)
    return idx, f.direct(den, idx, nrhs)


def test_laplace_uniform_65k_np4():
    c = bench.CONFIGS["laplace-uniform-65k"]
    seed = 0x4B1F0000 + c["cfg"]
    pos, nor = bench.make_points(c, seed)
    print(f"\n[case] laplace-uniform-65k n={c['n']} np={c['np']} ptsmax={c['ptsmax']}", flush=True)
    f = setup(c, pos, nor, c["ptsmax"])
    den = bench.make_density(c, 1, seed)[0]
    plain = f.eval(den).copy()
    mixed = f.eval(den, mode=K.MODE_MIXED).copy()
    o = O.OracleFMM(c["kt"], c["coefs"], c["np"], c["ptsmax"], 10, c["ctr"], c["rootlvl"])
    o.setup(pos, nor, pos, pinv=(f.operator(0), f.operator(1)))
    ref = o.eval(den)
    e = relerr(plain, ref)
    emax = float(np.max(np.abs(plain - ref)) / np.max(np.abs(ref)))
    em = relerr(mixed, plain)
    idx, d = check_vs_direct(f, den, 1, c["n"], 1, c["nchk"], seed)
    ep, epf = errors(plain[idx], d[0])
    ex, exf = errors(mixed[idx], d[0])
    print(f"  plain vs oracle {e:.3e} (max-abs/max {emax:.3e}); mixed vs plain {em:.3e}; vs direct on "
          f"{c['nchk']} targets: plain {ep:.3e} (FMM norm {epf:.3e}), mixed {ex:.3e} (FMM norm {exf:.3e})", flush=True)
    assert e <= PARITY and emax <= PARITY
    assert em <= MIXED_VS_PLAIN
    assert ep <= FMM_BAND[4] and ex <= FMM_BAND[4]


def test_navier_clusters_1m_np8():
    c = bench.CONFIGS["navier-clusters-1m"]
    seed = 0x4B1F0000 + c["cfg"]
    pos, nor = bench.make_points(c, seed)
    print(f"\n[case] navier-clusters-1m n={c['n']} np={c['np']} ptsmax={c['ptsmax']}", flush=True)
    t0 = time.time()
    f = setup(c, pos, nor, c["ptsmax"])
    den = bench.make_density(c, 3, seed)[0]
    plain = f.eval(den).copy()
    mixed = f.eval(den, mode=K.MODE_MIXED).copy()
    w = f.list_csr(2)[1].size
    x = f.list_csr(3)[1].size
    print(f"  GPU set-up + evaluates {time.time() - t0:.1f} s; W entries {w}, X entries {x}; near-field entries "
          f"left to the mixed form's FP64 pass {int(f.plan_stats()['near_flagged'])}", flush=True)
    assert w > 0 and x > 0, "the clustered workload must exercise the W and X lists"
    em = relerr(mixed, plain)
    idx, d = check_vs_direct(f, den, 1, c["n"], 3, c["nchk"], seed)
    ep, epf = errors(plain.reshape(-1, 3)[idx].ravel(), d[0])
    ex, _ = errors(mixed.reshape(-1, 3)[idx].ravel(), d[0])
    print(f"  mixed vs plain {em:.3e}; vs direct on {c['nchk']} targets: plain {ep:.3e} (FMM norm {epf:.3e}), "
          f"mixed {ex:.3e}", flush=True)
    t0 = time.time()
    o = O.OracleFMM(c["kt"], c["coefs"], c["np"], c["ptsmax"], 10, c["ctr"], c["rootlvl"])
    o.setup(pos, nor, pos, pinv=(f.operator(0), f.operator(1)))
    ref = o.eval(den)
    e = relerr(plain, ref)
    emax = float(np.max(np.abs(plain - ref)) / np.max(np.abs(ref)))
    print(f"  plain vs oracle on every target {e:.3e} (max-abs/max {emax:.3e}); oracle {time.time() - t0:.1f} s",
          flush=True)
    assert e <= PARITY_DEEP and emax <= PARITY
    assert em <= MIXED_VS_PLAIN
    assert ep <= FMM_BAND[8] and ex <= FMM_BAND[8]


def test_laplace_uniform_4m_16rhs():
    c = bench.CONFIGS["laplace-uniform-4m-16rhs"]
    seed = 0x4B1F0000 + c["cfg"]
    pos, nor = bench.make_points(c, seed)
    n, nrhs = c["n"], c["nrhs"]
    print(f"\n[case] laplace-uniform-4m-16rhs n={n} nrhs={nrhs} np=6 ptsmax=256", flush=True)
    den = bench.make_density(c, 1, seed)
    f = setup(c, pos, nor, 256)
    plain = f.eval(den, nrhs=nrhs).reshape(nrhs, n).copy()
    mixed = f.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, n).copy()
    em = max(relerr(mixed[r], plain[r]) for r in range(nrhs))
    # the wave-per-tile M2L on every level (at level 5 the FP64 spectra span
    # 8.5 GB, so its halo addresses take the 64-bit path) against the
    # block-per-tile default: the same sums in another order
    nlev = len(f.settings(K.MODE_PLAIN)["m2l"])
    f.set_settings(K.MODE_PLAIN, -1, [4] * nlev)
    f.set_settings(K.MODE_MIXED, -1, [4] * nlev)
    pw = f.eval(den, nrhs=nrhs).reshape(nrhs, n).copy()
    mw = f.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, n).copy()
    ewp = max(relerr(pw[r], plain[r]) for r in range(nrhs))
    ewm = max(relerr(mw[r], mixed[r]) for r in range(nrhs))
    print(f"  wave-per-tile M2L against block-per-tile: plain {ewp:.3e}, mixed {ewm:.3e}", flush=True)
    assert ewp <= 1e-13 and ewm <= 5e-7
    del pw, mw
    idx, d = check_vs_direct(f, den, nrhs, n, 1, c["nchk"], seed)
    ep, epf = errors(plain[:, idx], d)
    ex, exf = errors(mixed[:, idx], d)
    e0, _ = errors(plain[0, idx], d[0])
    print(f"  np 6: mixed vs plain {em:.3e}; vs direct on {c['nchk']} targets x {nrhs} RHS: plain {ep:.3e} "
          f"(FMM norm {epf:.3e}), mixed {ex:.3e} (FMM norm {exf:.3e}); first RHS alone: plain {e0:.3e}",
          flush=True)
    assert em <= MIXED_VS_PLAIN
    # the truncation error of np 6 on this workload: ~1.3e-6 in both forms
    assert 5e-7 <= ep <= 2e-6 and 5e-7 <= ex <= 2e-6
    assert abs(ex - ep) <= 0.1 * ep
    del f
    g = setup(c, pos, nor, 256, np_=8)
    p8 = g.eval(den, nrhs=nrhs).reshape(nrhs, n).copy()
    m8 = g.eval(den, nrhs=nrhs, mode=K.MODE_MIXED).reshape(nrhs, n).copy()
    e8, e8f = errors(p8[:, idx], d)
    x8, _ = errors(m8[:, idx], d)
    em8 = max(relerr(m8[r], p8[r]) for r in range(nrhs))
    print(f"  np 8: vs direct plain {e8:.3e} (FMM norm {e8f:.3e}), mixed {x8:.3e}; mixed vs plain {em8:.3e}",
          flush=True)
    assert e8 <= 1e-6 and x8 <= 1e-6 and em8 <= MIXED_VS_PLAIN


def test_batch_64_problems_mixed():
    nrhs = 8
    probs = bench.batch_problems()
    fmms, srcs, trgs, eager, dens = [], [], [], [], []
    t0 = time.time()
    for pr in probs:
        f = K.FMM3d("fmm3d_")
        f.srcpos, f.srcnor, f.trgpos = pr["pos"], pr["nor"], pr["pos"]
        f.ctr, f.rootlvl = pr["ctr"], pr["rootlvl"]
        f.knl = K.Kernel3d(pr["kt"], pr["coefs"])
        f.setup({"-fmm3d_np": str(pr["np"]), "-fmm3d_let3d_ptsmax": "100", "-fmm3d_let3d_maxlevel": "10"})
        den = bench.batch_density(pr, f.knl.sdof(), nrhs)
        s = K.DeviceBuffer(den.nbytes)
        s.upload(den)
        t = K.DeviceBuffer(nrhs * pr["n"] * f.knl.tdof() * 8)
        f.eval_device(s, t, nrhs, K.MODE_MIXED)
        out = np.empty((nrhs, pr["n"] * f.knl.tdof()))
        t.download(out)
        fmms.append(f); srcs.append(s); trgs.append(t); eager.append(out); dens.append(den)
    print(f"\n[case] batch-64 mixed: set-up and eager evaluates {time.time() - t0:.1f} s", flush=True)
    b = K.Batch(fmms)
    for rep in range(2):
        for t in trgs:
            t.upload(np.zeros(t.nbytes // 8))
        b.eval_device(srcs, trgs, [nrhs] * len(fmms), K.MODE_MIXED)
        K.synchronize()
        print(f"  batch launch {rep}: {b.time():.1f} ms", flush=True)
        for i, f in enumerate(fmms):
            got = np.empty_like(eager[i])
            trgs[i].download(got)
            assert np.array_equal(got, eager[i]), f"launch {rep}: problem {i} differs from its eager evaluate"
    worst = {}
    for i, (pr, f) in enumerate(zip(probs, fmms)):
        tdof = f.knl.tdof()
        idx = np.sort(np.random.default_rng(pr["seed"] + 2).choice(pr["n"], 200, replace=False)).astype(np.int32)
        d = f.direct(dens[i], idx, nrhs)
        v = eager[i].reshape(nrhs, pr["n"], tdof)[:, idx, :].reshape(nrhs, -1)
        e, _ = errors(v, d)
        key = (pr["kt"], pr["np"])
        worst[key] = max(worst.get(key, (0.0, -1)), (e, i))
        # the Stokes double layer on small spheres reaches ~2e-2 at np 4 and
        # ~3e-4 at np 6 (the algorithm's error on that geometry: see below)
        band = {(321, 4): 5e-2, (321, 6): 1e-3}.get(key, FMM_BAND[pr["np"]])
        assert e <= band, f"problem {i} {key}: error {e:.3e} against the direct sum"
    print("  worst error vs direct per (kernel, np): " +
          ", ".join(f"{k}: {v[0]:.3e} (problem {v[1]})" for k, v in sorted(worst.items())), flush=True)
    # the worst Stokes problem of each np: the plain GPU form equals the
    # oracle, and the oracle's own error against the direct sum is the same,
    # so the error is that of the algorithm on this geometry, not of this
    # implementation
    for np_ in (4, 6):
        if (321, np_) not in worst:
            continue
        e_s, i = worst[(321, np_)]
        pr, f = probs[i], fmms[i]
        plain = f.eval(dens[i][0]).copy()
        o = O.OracleFMM(321, pr["coefs"], pr["np"], 100, 10, pr["ctr"], pr["rootlvl"])
        o.setup(pr["pos"], pr["nor"], pr["pos"], pinv=(f.operator(0), f.operator(1)))
        ref = o.eval(dens[i][0])
        idx = np.sort(np.random.default_rng(pr["seed"] + 2).choice(pr["n"], 200, replace=False)).astype(np.int32)
        d = f.direct(dens[i][:1], idx, 1)[0]
        eo, _ = errors(ref.reshape(-1, 3)[idx].ravel(), d)
        eg, _ = errors(plain.reshape(-1, 3)[idx].ravel(), d)
        print(f"  problem {i} (Stokes DL, n={pr['n']}, np {np_}): plain vs oracle {relerr(plain, ref):.3e}; "
              f"vs direct, first RHS: oracle {eo:.3e}, GPU plain {eg:.3e}", flush=True)
        assert relerr(plain, ref) <= PARITY
        assert abs(eo - eg) <= 1e-6 * eo
    b.destroy()
