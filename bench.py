#!/usr/bin/env python3
"""Headline benchmark: DGEMM 8192^3 (NN, alpha=1, beta=0) on one MI355X.

One step = one C <- alpha*A*B + beta*C call of the library's native FP64
matrix-core DGEMM through the C ABI (device-pointer entry point), inputs
already resident in HBM (seeded splitmix64 uniform[-1,1) filled on the
device).  W untimed warm-up steps, then exactly K steps between two HIP
events on the launch stream, with a device synchronize on both sides.

Prints one JSON line.  `roofline` prices the dominant kernel against the
FP64 matrix-core peak (78.6 TFLOP/s); `cpu_baseline` times the reference's
own CPU dgemm_ (built from its sources into oracle/_ref) on this host's cores
on the same workload, when the host has AVX-512F.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_INT8_TOPS = 5000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "dgemm_8192_nn_traffic.json")


def _time(fn, stream, steps, warmup):
    from gemm_avx512f_amd.device import Event, synchronize
    for _ in range(warmup):
        fn()
    stream.synchronize()
    synchronize()
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    ms = e0.elapsed_ms(e1)
    synchronize()
    return ms, time.perf_counter() - t0


def _sampled_error(a, b, c, rows, cols):
    """Max and RMS relative error of C at sampled entries against long-double
    dot products of the same (downloaded) rows of A and columns of B."""
    ref = np.array([np.dot(a[i, :].astype(np.longdouble), b[:, j].astype(np.longdouble)) for i, j in zip(rows, cols)])
    got = c[rows, cols].astype(np.longdouble)
    scale = np.max(np.abs(ref))
    d = np.abs(got - ref)
    return float(np.max(d) / scale), float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(ref * ref)))


def gpu_leg(size, steps, warmup, variants=(("crt", 16), ("slices", 8)), seed=0):
    """Native FP64 DGEMM (timed region = the metric) and the emulated DGEMMs
    (int8 matrix cores: modular/CRT form and slice form) on the same inputs,
    each with its sampled error."""
    from gemm_avx512f_amd import _lib, blas
    from gemm_avx512f_amd.device import DeviceMatrix, Stream, fill_uniform

    L = _lib.lib()
    _lib.check(L.gemm_amd_init(), "init")
    m = n = k = size
    stream = Stream()
    A = DeviceMatrix(m, k, np.float64)
    B = DeviceMatrix(k, n, np.float64)
    C = DeviceMatrix(m, n, np.float64)
    d = fill_uniform(A, seed, 0, stream=stream)
    d = fill_uniform(B, seed, d, stream=stream)
    fill_uniform(C, seed, d, stream=stream)
    out = {}
    ms, wall = _time(lambda: blas.dgemm("N", "N", m, n, k, 1.0, A, B, 0.0, C, stream=stream), stream, steps, warmup)
    out["native"] = {"ms": ms, "wall": wall}
    rng = np.random.default_rng(12345)
    rows, cols = rng.integers(0, m, 256), rng.integers(0, n, 256)
    a, b = A.to_host(), B.to_host()
    c_nat = C.to_host()
    if not np.isfinite(c_nat).all():
        raise RuntimeError("non-finite DGEMM result")
    out["native"]["err"] = _sampled_error(a, b, c_nat, rows, cols)
    del c_nat
    out["emulated"] = []
    for kind, count in variants:
        C2 = DeviceMatrix(m, n, np.float64)
        if kind == "crt":
            fn = lambda: blas.dgemm_ozaki_crt("N", "N", m, n, k, 1.0, A, B, 0.0, C2, nmoduli=count,  # noqa: E731
                                              stream=stream)
            products = count
            path = "gemm_amd_dgemm_ozaki_crt (int8 residues, %d moduli, %d int8 GEMMs)" % (count, count)
        else:
            fn = lambda: blas.dgemm_ozaki("N", "N", m, n, k, 1.0, A, B, 0.0, C2, nslices=count,  # noqa: E731
                                          stream=stream)
            products = count * (count + 1) // 2
            path = "gemm_amd_dgemm_ozaki (int8 slices, %d slices, %d int8 GEMMs)" % (count, products)
        ms2, wall2 = _time(fn, stream, steps, warmup)
        out["emulated"].append({"kind": kind, "count": count, "products": products, "path": path, "ms": ms2,
                                "wall": wall2, "err": _sampled_error(a, b, C2.to_host(), rows, cols)})
        del C2
    return out


def cpu_leg(size, reps=3):
    """Reference dgemm_ (oracle/_ref/DGEMM.so, -O2 -mavx512f -fopenmp) on this host."""
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    threads = int(os.environ["OMP_NUM_THREADS"])
    kind = "reference"
    if O.ref_lib("DGEMM.so") is None:
        kind = "port"
    m = n = k = size
    s = O.Stream(0)
    a, b, c = s.f64(m, k), s.f64(k, n), s.f64(m, n)
    best = None
    for i in range(reps + 1):
        t0 = time.perf_counter()
        if kind == "reference":
            O.ref_gemm(np.float64, "N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
        else:
            O.dgemm_blocked("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, threads)
        dt = time.perf_counter() - t0
        if i > 0:
            best = dt if best is None else min(best, dt)
    tflops = 2.0 * m * n * k / best / 1e12
    return {"value": round(tflops, 4), "unit": "TFLOP/s", "cores": threads, "kind": kind,
            "sample": "DGEMM %d^3 NN alpha=1 beta=0, best of %d after 1 warm-up (%.2f s per call)" % (size, reps, best)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-emulated", action="store_true", help="skip the emulated (int8) DGEMM legs")
    ap.add_argument("--nslices", type=int, default=8, help="slices per operand of the slice-form emulated DGEMM")
    ap.add_argument("--nmoduli", type=int, default=16, help="moduli of the modular-form emulated DGEMM")
    ap.add_argument("--cpu-size", type=int, default=None, help="CPU leg problem size (default: --size)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB first (a known-size kernel that "
                         "calibrates the HBM byte counters when the run is profiled)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the DGEMM on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    size = args.size
    if args.copy_mb > 0:
        from gemm_avx512f_amd.device import DeviceBuffer, synchronize
        from gemm_avx512f_amd import _lib
        nb = args.copy_mb << 20
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        _lib.check(_lib.lib().gemm_amd_memcpy(dst.ptr, src.ptr, nb, 3), "calibration copy")
        synchronize()
        del src, dst
    variants = () if args.no_emulated else (("crt", args.nmoduli), ("slices", args.nslices))
    legs = gpu_leg(size, args.steps, args.warmup, variants=variants)
    ms, wall = legs["native"]["ms"], legs["native"]["wall"]
    flops = 2.0 * size ** 3
    ms_step = ms / args.steps
    tflops = flops * args.steps / (ms / 1e3) / 1e12
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            prof = json.load(f)
        if prof.get("size") == size:
            traffic = prof.get("hbm_bytes_per_launch")
    result = {
        "metric": "TFLOP/s & %peak: DGEMM 8192^3 native FP64 vs emulated at FP64-level error",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: splitmix64 uniform[-1,1) (seed 0) filled in HBM",
        "config": {"workload": "DGEMM %d^3 NN, alpha=1, beta=0, column-major, native FP64 matrix cores" % size,
                   "m": size, "n": size, "k": size, "path": "gemm_amd_dgemm (device pointers)"},
        "pct_peak": round(100.0 * tflops / PEAK_FP64_TFLOPS, 2),
        "roofline": {"bound": "mfma", "achieved": round(flops / (ms_step / 1e3) / 1e12, 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / (ms_step / 1e3) / 1e12 / PEAK_FP64_TFLOPS, 4), "traffic": traffic},
        "native_err": {"max_rel": legs["native"]["err"][0], "rms_rel": legs["native"]["err"][1],
                       "sample": "256 entries vs long-double dot products"},
        "wall_s_timed_region": round(wall, 3),
    }
    variants = []
    for e in legs["emulated"]:
        em_tf = flops * args.steps / (e["ms"] / 1e3) / 1e12
        variants.append({
            "path": e["path"],
            "tflops_fp64_equiv": round(em_tf, 3),
            "pct_fp64_peak": round(100.0 * em_tf / PEAK_FP64_TFLOPS, 2),
            "int8_tops": round(em_tf * e["products"], 1),
            "int8_frac_of_peak": round(em_tf * e["products"] / PEAK_INT8_TOPS, 4),
            "ms_per_step": round(e["ms"] / args.steps, 4),
            "max_rel": e["err"][0], "rms_rel": e["err"][1],
            "fp64_level": bool(e["err"][0] <= legs["native"]["err"][0]),
            "faster_than_native": bool(em_tf > tflops),
        })
    if variants:
        # headline emulated figure: the fastest variant whose error is at FP64 level
        ok = [v for v in variants if v["fp64_level"]] or variants
        result["emulated"] = max(ok, key=lambda v: v["tflops_fp64_equiv"])
        result["emulated_variants"] = variants
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_leg(args.cpu_size or size)
        except Exception as exc:  # reported, never silently replaced
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
