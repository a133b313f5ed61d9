#!/usr/bin/env python3
"""Evaluated targets/sec of the KIFMM3d evaluate pass on one MI355X.

A step is one evaluate pass (all right-hand sides) over device-resident
densities: warm-up steps first, then K timed steps between device
synchronisations.  The default workload is Laplace single layer,
N = 4,194,304 points uniform in [0,1)^3, 16 density vectors, np = 6.  The
error of the last step is checked against the FP64 direct sum (on the GPU) at
sampled targets, and the CPU oracle is timed on a bounded sample of the same
workload for the CPU baseline.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "kernel-independent-fast-multipole-method_amd")
sys.path.insert(0, PKG)

METRIC = "evaluated targets/sec at rel err ≤1e-6 vs FP64 direct (chosen vs plain FP64 GPU)"
FP64_PEAK_TF = 78.6      # MI355X FP64 (vector and matrix) TFLOP/s
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E GB/s

# seeds 0x4B1F0000 + config number
CONFIGS = {
    "laplace-uniform-65k": dict(cfg=1, kt=111, coefs=[1.0], n=65536, np=4, ptsmax=100, nrhs=1, gen="cube",
                                ctr=(0.5, 0.5, 0.5), rootlvl=1, nchk=200),
    "stokes-dl-sphere-262k": dict(cfg=2, kt=321, coefs=[1.0], n=262144, np=6, ptsmax=100, nrhs=1, gen="sphere",
                                  ctr=(0.0, 0.0, 0.0), rootlvl=0, nchk=200),
    "navier-clusters-1m": dict(cfg=3, kt=511, coefs=[1.0, 0.3], n=1048576, np=8, ptsmax=150, nrhs=1,
                               gen="clusters32", ctr=(0.5, 0.5, 0.5), rootlvl=1, nchk=200),
    "laplace-uniform-4m-16rhs": dict(cfg=4, kt=111, coefs=[1.0], n=4194304, np=6, ptsmax=128, nrhs=16, gen="cube",
                                     ctr=(0.5, 0.5, 0.5), rootlvl=1, nchk=1000),
}
DEFAULT = "laplace-uniform-4m-16rhs"

# flops per source-target pair for nrhs right-hand sides (FMA = 2, sqrt/div = 1)
PAIR_FLOPS = {111: lambda r: 9 + 2 * r, 321: lambda r: 18 + 12 * r, 511: lambda r: 12 + 19 * r,
              301: lambda r: 14 + 20 * r}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_points(c, seed):
    rng = np.random.default_rng(seed)
    n = c["n"]
    if c["gen"] == "cube":
        p = rng.random((n, 3))
        return p, p
    if c["gen"] == "sphere":
        ct = rng.uniform(-1, 1, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        st = np.sqrt(1 - ct * ct)
        p = np.stack([st * np.cos(ph), st * np.sin(ph), ct], axis=1)
        return p, p.copy()
    if c["gen"] == "clusters32":
        k = 32
        ctr = rng.uniform(0.1, 0.9, (k, 3))
        sig = np.exp(rng.uniform(np.log(0.01), np.log(0.1), k))
        out = np.empty((0, 3))
        while out.shape[0] < n:
            lab = rng.integers(0, k, n)
            q = ctr[lab] + sig[lab, None] * rng.standard_normal((n, 3))
            out = np.concatenate([out, q[np.all((q >= 0) & (q < 1), axis=1)]])
        p = out[:n]
        return p, p
    raise ValueError(c["gen"])


def make_density(c, sdof, seed):
    rng = np.random.default_rng(seed + 1)
    if c["kt"] == 511:
        return rng.standard_normal((c["nrhs"], c["n"] * sdof))
    return rng.uniform(-1, 1, (c["nrhs"], c["n"] * sdof))


def cpu_baseline(c, seconds_target=15.0):
    """Time the CPU oracle (plain FP64 restatement of the reference's eval, one
    thread) on a bounded sample of the workload: the same generator, kernel, np
    and ptsmax with fewer points, one RHS.  Returns (targets/s, sample text)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes

    import oracle as O

    lib = O.lib()
    lib.orc_set_accurate.argtypes = [ctypes.c_int]
    lib.orc_set_accurate(0)  # plain FP64 accumulation, as the reference's BLAS does
    n = {111: 131072, 321: 32768, 511: 16384}[c["kt"]]
    s = dict(c, n=n, nrhs=1)
    pos, nor = make_points(s, 0x4B1F0000 + c["cfg"])
    o = O.OracleFMM(c["kt"], c["coefs"], c["np"], c["ptsmax"], 10, c["ctr"], c["rootlvl"])
    o.setup(pos, nor, pos)
    den = make_density(s, O.kernel_dofs(c["kt"])[0], 0)[0]
    o.eval(den)  # first eval builds the lazy operators, as in the reference
    t0 = time.perf_counter()
    o.eval(den)
    dt = time.perf_counter() - t0
    lib.orc_set_accurate(1)
    return n / dt, f"{c['gen']} N={n}, np={c['np']}, ptsmax={c['ptsmax']}, 1 RHS, warm eval {dt:.2f} s"


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=DEFAULT, choices=sorted(CONFIGS))
    ap.add_argument("--ptsmax", type=int, default=None)
    ap.add_argument("--nrhs", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-check", action="store_true", help="skip the direct-sum error check")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and not built in this version", file=sys.stderr)
        sys.exit(2)

    import kifmm3d as K

    c = dict(CONFIGS[args.config])
    if args.ptsmax:
        c["ptsmax"] = args.ptsmax
    if args.nrhs:
        c["nrhs"] = args.nrhs
    if K.device_count() < 1:
        raise SystemExit("no HIP device visible")
    seed = 0x4B1F0000 + c["cfg"]
    pos, nor = make_points(c, seed)
    f = K.FMM3d("fmm3d_")
    f.srcpos, f.srcnor, f.trgpos = pos, nor, pos
    f.ctr, f.rootlvl = c["ctr"], c["rootlvl"]
    f.knl = K.Kernel3d(c["kt"], c["coefs"])
    t0 = time.perf_counter()
    f.setup({"-fmm3d_np": str(c["np"]), "-fmm3d_let3d_ptsmax": str(c["ptsmax"]), "-fmm3d_let3d_maxlevel": "10"})
    t_setup = time.perf_counter() - t0
    info, stats = f.info(), f.plan_stats()
    log(f"{args.config}: setup {t_setup:.2f} s, {info['nnode']} nodes, {info['nlevel']} levels, "
        f"near pairs {stats['near_pairs']:.3e}, V pairs {stats['v_pairs']:.3e}")
    sdof, tdof, nrhs, n = f.knl.sdof(), f.knl.tdof(), c["nrhs"], c["n"]
    den = make_density(c, sdof, seed)
    d_src = K.DeviceBuffer(den.nbytes)
    d_src.upload(den)
    d_trg = K.DeviceBuffer(nrhs * n * tdof * 8)

    for i in range(args.warmup):
        f.eval_device(d_src, d_trg, nrhs)
        log(f"warmup {i + 1}/{args.warmup}: {f.stage_times()['total']:.1f} ms")
    stage_sum = {}
    K.synchronize()
    # (the per-kernel event timings read after the loop are those of the last timed step)
    t0 = time.perf_counter()
    for i in range(args.steps):
        f.eval_device(d_src, d_trg, nrhs)
        for k, v in f.stage_times().items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    K.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1000.0 * dt / args.steps
    stages = {k: v / args.steps for k, v in stage_sum.items()}
    value = n * nrhs / (dt / args.steps)
    log(f"timed {args.steps} steps: {ms_step:.2f} ms/step, {value:.4e} targets/s, stages {stages}")

    err = None
    if not args.no_check:
        val = np.empty((nrhs, n * tdof))
        d_trg.download(val)
        rng = np.random.default_rng(seed + 2)
        idx = np.sort(rng.choice(n, c["nchk"], replace=False)).astype(np.int32)
        d = f.direct(d_src, idx, nrhs)
        v = val.reshape(nrhs, n, tdof)[:, idx, :].reshape(nrhs, -1)
        err = float(np.linalg.norm(v - d) / np.linalg.norm(d))
        log(f"rel L2 error vs FP64 direct sum on {c['nchk']} targets x {nrhs} RHS: {err:.3e}")

    # roofline of the dominant kernel class, timed live with HIP events on the
    # library's stream during the last timed step
    kt_mm = {111: 111, 321: 301, 511: 511}[c["kt"]]
    sm, tm = {111: (1, 1), 301: (4, 3), 511: (3, 3)}[kt_mm]
    kt_times = f.kernel_times()
    nsp = {4: 56, 6: 152, 8: 296}[c["np"]]
    work = {  # (algorithmic amount per step, unit, bound)
        "p2p": (stats["near_pairs"] * PAIR_FLOPS[c["kt"]](nrhs), "flop", "mfma"),
        "s2m": (stats["s2m_pairs"] * PAIR_FLOPS[c["kt"]](nrhs), "flop", "mfma"),
        "s2l": (stats["s2l_pairs"] * PAIR_FLOPS[c["kt"]](nrhs), "flop", "mfma"),
        "s2p": (stats["far_pairs"] * PAIR_FLOPS[kt_mm](nrhs), "flop", "mfma"),
        "m2l": (stats["v_pairs"] * nrhs * stats["C"] * 8.0 * sm * tm, "flop", "mfma"),
        "gemm_up": (stats["up_flops"] * nrhs, "flop", "mfma"),
        "gemm_down": (stats["down_flops"] * nrhs, "flop", "mfma"),
        "fft_fwd": (stats["fft_fwd"] * nrhs * (nsp * 8.0 + stats["C"] * 16.0), "byte", "hbm"),
        "fft_inv": (stats["fft_inv"] * nrhs * (nsp * 8.0 + stats["C"] * 16.0), "byte", "hbm"),
    }
    rates = {}
    for k, (amt, unit, bound) in work.items():
        ms, calls = kt_times[k]
        if ms > 0 and calls > 0:
            rate = amt / (ms * 1e-3) / (1e12 if unit == "flop" else 1e9)
            rates[k] = (ms, calls, rate, unit, bound)
            log(f"kernel {k:9s}: {ms:9.3f} ms in {calls:3d} launch groups, "
                f"{rate:9.3f} {'TFLOP/s' if unit == 'flop' else 'GB/s'}")
    dom = max(rates, key=lambda k: rates[k][0])
    ms, calls, rate, unit, bound = rates[dom]
    peak = FP64_PEAK_TF if unit == "flop" else HBM_PEAK_GBS
    roofline = {"kernel": dom, "bound": bound, "achieved": round(rate, 4), "peak": peak,
                "unit": "TFLOP/s" if unit == "flop" else "GB/s", "frac": round(rate / peak, 5),
                "traffic": None, "ms_per_step": round(ms, 3), "launch_groups": calls}

    cpu = None
    if not args.no_cpu:
        cv, sample = cpu_baseline(c)
        cpu = {"value": cv, "unit": "targets/s", "cores": 1, "kind": "port", "sample": sample}
        log(f"cpu baseline: {cv:.4e} targets/s ({sample})")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.config, "form": "plain FP64", "kernel": c["kt"], "n_points": n, "nrhs": nrhs,
                   "np": c["np"], "ptsmax": c["ptsmax"], "rel_err_vs_direct": err, "setup_s": round(t_setup, 3),
                   "stage_ms": {k: round(v, 3) for k, v in stages.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
