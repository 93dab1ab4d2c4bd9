#!/usr/bin/env python3
"""Headline benchmark: frequency-point solves/s of the omegaFQFmu
frequency-dependent response on one MI355X.

Workload (the largest single-GPU configuration): omegaFQFmu Ag fcc sphere,
a = 4.09 A, r = 4 nm (15683 atoms, 62732 complex unknowns), Etchegoin
interband alpha(omega), 201 frequencies in 2.8-4.2 eV, 3 field directions,
GMRES on the fly.  One step solves the whole spectrum (201 x 3 systems) with
FP32 interaction tiles inside FP64 iterative refinement, many systems
sharing every operator product; the stopping test is the FP64 true residual
||b - A x|| < 1e-12 sqrt(n_var).

Reported beside it: the error of the polarizability against an FP64 solve
(tighter tolerance) at sampled frequencies, the same sample solved by the
plain FP64 GPU path (one system at a time, the reference's GMRES), the
roofline of the pair kernel (HIP events on the library's stream), and the
FP64 CPU oracle timed on this host.

    python bench.py [--gpus 1] [--steps 1] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense fp16 MFMA peak (the pair kernel's arithmetic unit)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--freqs", type=int, default=201)
    ap.add_argument("--radius", type=float, default=40.0)
    ap.add_argument("--batch", type=int, default=48)
    ap.add_argument("--krylov", type=int, default=120)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--check", type=int, default=3, help="frequencies checked against FP64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--traffic", action="store_true",
                    help="run one 48-system pair-kernel launch and a 256 MiB device copy (for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE), then exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    from plasmonx_amd import Context, _lib
    from plasmonx_amd.workloads import c2_silver_sphere

    t0 = time.time()
    s, freq = c2_silver_sphere(radius=a.radius, n_freq=a.freqs)
    aw = s.alpha_table(freq)
    e0 = 1.0e-4
    tol = a.tol * np.sqrt(s.n_var)
    log(f"[bench] C2 omegaFQFmu Ag sphere: {s.n_atoms} atoms, {s.n_var} unknowns, {len(freq)} freq x 3 fields "
        f"(setup {time.time() - t0:.1f}s)")
    ctx = Context(0)
    ctx.set_atoms(s)
    ctx.set_alpha(aw)
    ctx.set_option("batch", a.batch)
    if a.traffic:
        traffic_run(ctx, s, freq)
        return

    def solve(fr_idx, precision=_lib.MIXED, tol_abs=tol, pol=7, kdim=a.krylov):
        ctx.set_alpha(aw[:, fr_idx])
        return ctx.solve_spectrum(s.n_var, freq[fr_idx], e0, pol_mask=pol, method=_lib.ITERATIVE_OTF,
                                  tol_abs=tol_abs, krylov_dim=kdim, precision=precision)

    all_idx = np.arange(len(freq))
    for w in range(a.warmup):
        r = solve(all_idx[:4])
        log(f"[bench] warmup {w}: 4 freq ok, iters {r['iters'][:, 0].tolist()}")
    ctx.set_alpha(aw)
    times, stats = [], []
    for k in range(a.steps):
        t1 = time.perf_counter()
        out = ctx.solve_spectrum(s.n_var, freq, e0, pol_mask=7, method=_lib.ITERATIVE_OTF, tol_abs=tol,
                                 krylov_dim=a.krylov, precision=_lib.MIXED)
        t2 = time.perf_counter()
        times.append(t2 - t1)
        stats.append(ctx.stats())
        log(f"[bench] step {k}: {t2 - t1:.2f}s, {len(freq) / (t2 - t1):.2f} freq-points/s, "
            f"iters mean {out['iters'].mean():.1f}, max true residual {out['true_residual'].max():.2e} (tol {tol:.2e})")
    total = sum(times)
    value = a.steps * len(freq) / total
    st = stats[-1]
    assert np.all(out["true_residual"] < tol)

    # accuracy against FP64 at sampled frequencies (tighter tolerance)
    chk = np.unique(np.linspace(0, len(freq) - 1, max(1, a.check)).astype(int))
    ctx.set_option("batch", 1)
    tp = time.perf_counter()
    ref = solve(chk, precision=_lib.FP64, tol_abs=1e-2 * tol, kdim=600)
    t_ref = time.perf_counter() - tp
    errs = [float(np.linalg.norm(out["polar"][f] - ref["polar"][k]) / np.linalg.norm(ref["polar"][k]))
            for k, f in enumerate(chk)]
    log(f"[bench] max rel polar error vs FP64 at {len(chk)} freq: {max(errs):.2e}")

    plain = None
    if not a.no_plain:
        tp = time.perf_counter()
        pr = solve(chk, precision=_lib.FP64, tol_abs=tol, kdim=600)
        tpl = time.perf_counter() - tp
        plain = {"value": len(chk) / tpl, "sample": f"{len(chk)} freq x 3 fields, one system at a time, "
                                                       f"reference GMRES in FP64",
                 "iters_mean": float(pr["iters"].mean())}
        log(f"[bench] plain FP64 GPU path: {plain['value']:.3f} freq-points/s ({tpl:.1f}s), iters {pr['iters'].tolist()}")

    # roofline of the dominant kernel: the MFMA pair sweep.  Algorithmic flops
    # per launch: the real 4N x 4N T applied to 2 real columns per system (2 flops
    # per element and column, 64 N(N-1) per system) plus 42 flops per ordered
    # pair for the elements; divided by the kernel's own time (HIP events on
    # the library's stream around each launch, summed over the timed step).
    ms = st["mfma_kernel_ms"]
    launches = st["mfma_launches"]
    achieved = st["sweep_flops_fp32"] / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": FP16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP16_MFMA_PEAK_TFLOPS, 4), "traffic": TRAFFIC_BYTES_PER_LAUNCH,
            "kernel": "k_tsweep_mfma<3> (fp16 hi/lo split, 3 MFMA products per element pair)",
            "launches": int(launches), "avg_launch_ms": round(ms / max(launches, 1), 3)}
    latency = {"p50_ms": round(st["latency_p50_ms"], 1), "p99_ms": round(st["latency_p99_ms"], 1),
               "definition": "per frequency point: admission of its first field to convergence of its last, "
                             "201 points queued at t=0 into 48 solver slots"}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(s, freq, aw, plain["iters_mean"] if plain else out["iters"].mean())
        log(f"[bench] cpu oracle: {cpu}")

    line = {
        "metric": "freq-point solves/s at <=1e-6 rel err vs FP64",
        "value": round(value, 4), "unit": "freq-point solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 tiles + fp64 refinement", "data": "synthetic",
        "config": {"workload": "omegaFQFmu Ag fcc sphere a=4.09A r=4nm, Etchegoin interband alpha, "
                               "201 w 2.8-4.2 eV x 3 fields, GMRES on the fly",
                   "n_atoms": s.n_atoms, "n_var": s.n_var, "n_freq": len(freq), "fields": 3,
                   "tol_abs": tol, "batch": a.batch, "krylov_dim": a.krylov, "parallelism": "single GPU"},
        "max_rel_err_vs_fp64": max(errs),
        "plain_fp64_gpu": plain,
        "speedup_vs_plain_fp64": round(value / plain["value"], 2) if plain else None,
        "iters_mean": float(out["iters"].mean()),
        "latency": latency,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


# HBM bytes per launch of the pair kernel (FETCH_SIZE + WRITE_SIZE from rocprofv3
# counters on the 48-system product, scaled by the same counters on a copy of
# known size); None until measured.
TRAFFIC_BYTES_PER_LAUNCH = 697.8e6


def traffic_run(ctx, s, freq):
    """Calibration copy of known size (a 256 MiB device-to-device hipMemcpy:
    reads and writes 256 MiB), then one mixed product of 48 systems: exactly
    one launch of the MFMA pair kernel."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")  # the HIP runtime libpxh is linked against
    size = 256 * 1024 * 1024
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(size)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemcpy(dst, src, ctypes.c_size_t(size), 3) != 0 or hip.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipMemcpy failed")
    hip.hipFree(src)
    hip.hipFree(dst)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(48, s.n_var)) + 1j * rng.normal(size=(48, s.n_var))
    ctx.set_option("matvec_mixed", 1)
    ctx.matvec(freq[len(freq) // 2], len(freq) // 2, x)
    ctx.set_option("matvec_mixed", 0)
    log("[bench] traffic run: 256 MiB copy + one 48-system product done")


def cpu_baseline(s, freq, aw, iters_per_system):
    """The FP64 oracle's on-the-fly product on this host (OpenMP), one
    product at full size; a frequency point costs iters x 3 products."""
    from oracle import pxo
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    pxo.load()
    rng = np.random.default_rng(0)
    x = rng.normal(size=s.n_var) + 1j * rng.normal(size=s.n_var)
    t = time.perf_counter()
    pxo.matvec_otf(s, freq[0], aw[:, 0], x)
    dt = time.perf_counter() - t
    per_point = dt * iters_per_system * 3
    return {"value": round(1.0 / per_point, 6), "unit": "freq-point solves/s", "cores": cores, "kind": "port",
            "sample": f"one FP64 on-the-fly product at full size ({dt:.1f}s) x {iters_per_system:.0f} GMRES "
                      f"iterations x 3 fields per frequency point"}


if __name__ == "__main__":
    main()
