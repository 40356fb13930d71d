#!/usr/bin/env python3
"""Headline benchmark: DGEMM 8192^3 (NN, alpha=1, beta=0) on one MI355X.

One step = one C <- alpha*A*B + beta*C call of the library's native FP64
matrix-core DGEMM through the C ABI (device-pointer entry point), inputs
already resident in HBM (seeded splitmix64 uniform[-1,1) filled on the
device).  W untimed warm-up steps, then exactly K steps between two HIP
events on the launch stream, with a device synchronize on both sides.

Prints one JSON line.  `roofline` prices the dominant kernel against the
FP64 matrix-core peak (78.6 TFLOP/s); `emulated` is the fastest int8
emulation whose sampled max error is no larger than native FP64's on the
headline inputs and on inputs of every dynamic range r in {0, 8, 32};
`dropin_host_pointer` is the reference's own entry point dgemm_ on host
arrays, PCIe copies included; `cpu_baseline` times the reference's own CPU
dgemm_ (built from its sources into oracle/_ref) on all physical cores of
this host, one OpenMP thread per core with close binding, on the same
workload, when the host has AVX-512F.
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


def _progress(msg):
    """One line per leg on stderr (the JSON result is the only line on stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


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


SAMPLES = 1024   # sampled entries of C per error figure


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
    rows, cols = rng.integers(0, m, SAMPLES), rng.integers(0, n, SAMPLES)
    a, b = A.to_host(), B.to_host()
    c_nat = C.to_host()
    if not np.isfinite(c_nat).all():
        raise RuntimeError("non-finite DGEMM result")
    out["native"]["err"] = _sampled_error(a, b, c_nat, rows, cols)
    out["host_ab"] = (a, b)
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
                                "accuracy_key": ("crt_%d_moduli" if kind == "crt" else "slices_%d") % count,
                                "wall": wall2, "err": _sampled_error(a, b, C2.to_host(), rows, cols)})
        del C2
    return out


def accuracy_leg(size, ranges=(0, 8, 32), seed=7):
    """Error vs exact (long-double) dot products at sampled entries of C for
    A, B entries u * 2^e, e uniform in [-r, r]: native DGEMM, the CRT form at
    12..16 moduli and the slice form at 4..8 slices (one call each)."""
    from gemm_avx512f_amd import blas
    from gemm_avx512f_amd.device import FILL_F64_RANGE, DeviceMatrix, Stream, fill_uniform

    m = n = k = size
    stream = Stream()
    A, B, C = DeviceMatrix(m, k, np.float64), DeviceMatrix(k, n, np.float64), DeviceMatrix(m, n, np.float64)
    rng = np.random.default_rng(999)
    rows, cols = rng.integers(0, m, SAMPLES), rng.integers(0, n, SAMPLES)
    out = {"workload": "DGEMM %d^3 NN, entries u*2^e, e uniform in [-r, r]; %d sampled entries" % (size, SAMPLES)}
    runs = [("native_fp64", lambda: blas.dgemm("N", "N", m, n, k, 1.0, A, B, 0.0, C, stream=stream))]
    for nm in (16, 15, 14, 13, 12):
        runs.append(("crt_%d_moduli" % nm, lambda nm=nm: blas.dgemm_ozaki_crt("N", "N", m, n, k, 1.0, A, B, 0.0, C,
                                                                            nmoduli=nm, stream=stream)))
    for ns in (8, 7, 6, 5, 4):
        runs.append(("slices_%d" % ns, lambda ns=ns: blas.dgemm_ozaki("N", "N", m, n, k, 1.0, A, B, 0.0, C,
                                                                     nslices=ns, stream=stream)))
    for r in ranges:
        d = fill_uniform(A, seed + r, 0, kind=FILL_F64_RANGE, range_r=r, stream=stream)
        fill_uniform(B, seed + r, d, kind=FILL_F64_RANGE, range_r=r, stream=stream)
        a, b = A.to_host(), B.to_host()
        res = {}
        for name, fn in runs:
            fn()
            stream.synchronize()
            e = _sampled_error(a, b, C.to_host(), rows, cols)
            res[name] = {"max_rel": e[0], "rms_rel": e[1]}
        out["r=%d" % r] = res
    return out


def sgemm_leg(size, steps, warmup, seed=0):
    """SGEMM size^3 NN: native FP32 matrix cores vs the bf16x3 / bf16x6 split
    forms on the same FP32 inputs, each with its sampled error."""
    from gemm_avx512f_amd import blas
    from gemm_avx512f_amd.device import DeviceMatrix, Stream, fill_uniform

    m = n = k = size
    stream = Stream()
    A, B, C = DeviceMatrix(m, k, np.float32), DeviceMatrix(k, n, np.float32), DeviceMatrix(m, n, np.float32)
    d = fill_uniform(A, seed, 0, stream=stream)
    fill_uniform(B, seed, d, stream=stream)
    a, b = A.to_host().astype(np.float64), B.to_host().astype(np.float64)
    rng = np.random.default_rng(54321)
    rows, cols = rng.integers(0, m, SAMPLES), rng.integers(0, n, SAMPLES)
    flops = 2.0 * m * n * k
    out = {"workload": "SGEMM %d^3 NN, alpha=1, beta=0, uniform[-1,1) FP32" % size}
    runs = (("native_fp32", lambda: blas.sgemm("N", "N", m, n, k, 1.0, A, B, 0.0, C, stream=stream), 157.3),
            ("bf16x3", lambda: blas.sgemm_bf16x("N", "N", m, n, k, 1.0, A, B, 0.0, C, nprod=3, stream=stream), 157.3),
            ("bf16x6", lambda: blas.sgemm_bf16x("N", "N", m, n, k, 1.0, A, B, 0.0, C, nprod=6, stream=stream), 157.3))
    for name, fn, peak in runs:
        ms, _ = _time(fn, stream, steps, warmup)
        tf = flops * steps / (ms / 1e3) / 1e12
        err = _sampled_error(a, b, C.to_host().astype(np.float64), rows, cols)
        out[name] = {"tflops": round(tf, 2), "pct_fp32_peak": round(100.0 * tf / peak, 2),
                     "ms_per_step": round(ms / steps, 4), "max_rel": err[0], "rms_rel": err[1]}
    return out


def configs_leg(steps=20, warmup=3, seed=0):
    """The smaller workloads next to the headline: SGEMM 1024^3 NN (alpha 1,
    beta 0); DGEMM 4096^3 NN and TN with alpha 1.5, beta 0.5; DTRSM side L,
    lower, no transpose, non-unit, M=4096, N=512, triangle uniform[-1,1),
    diagonal 4096 + uniform[0,1).  Each with its sampled error (GEMMs) or
    residual ||A X - B|| / (||A|| ||X|| M) (DTRSM)."""
    from gemm_avx512f_amd import blas
    from gemm_avx512f_amd.device import DeviceMatrix, Stream, fill_uniform

    stream = Stream()
    rng = np.random.default_rng(2468)
    out = {}
    n = 1024
    A, B, C = (DeviceMatrix(n, n, np.float32) for _ in range(3))
    d = fill_uniform(A, seed, 0, stream=stream)
    fill_uniform(B, seed, d, stream=stream)
    ms, _ = _time(lambda: blas.sgemm("N", "N", n, n, n, 1.0, A, B, 0.0, C, stream=stream), stream, steps, warmup)
    rows, cols = rng.integers(0, n, SAMPLES), rng.integers(0, n, SAMPLES)
    err = _sampled_error(A.to_host().astype(np.float64), B.to_host().astype(np.float64),
                         C.to_host().astype(np.float64), rows, cols)
    out["sgemm_1024_nn"] = {"tflops": round(2.0 * n ** 3 * steps / (ms / 1e3) / 1e12, 2),
                            "ms_per_call": round(ms / steps, 4), "max_rel": err[0], "rms_rel": err[1]}
    n = 4096
    A, B, C = (DeviceMatrix(n, n, np.float64) for _ in range(3))
    d = fill_uniform(A, seed, 0, stream=stream)
    d = fill_uniform(B, seed, d, stream=stream)
    fill_uniform(C, seed, d, stream=stream)
    a, b, c0 = A.to_host(), B.to_host(), C.to_host()
    rows, cols = rng.integers(0, n, SAMPLES), rng.integers(0, n, SAMPLES)
    for ta in ("N", "T"):
        C.buf.upload(np.asfortranarray(c0))
        blas.dgemm(ta, "N", n, n, n, 1.5, A, B, 0.5, C, stream=stream)   # one checked call from the known C
        stream.synchronize()
        opa = a if ta == "N" else a.T
        ref = np.array([1.5 * np.dot(opa[i, :].astype(np.longdouble), b[:, j].astype(np.longdouble))
                        + 0.5 * np.longdouble(c0[i, j]) for i, j in zip(rows, cols)])
        got = C.to_host()[rows, cols].astype(np.longdouble)
        ms, _ = _time(lambda: blas.dgemm(ta, "N", n, n, n, 1.5, A, B, 0.5, C, stream=stream), stream, steps, warmup)
        out["dgemm_4096_%sn" % ta.lower()] = {
            "tflops": round(2.0 * n ** 3 * steps / (ms / 1e3) / 1e12, 2), "ms_per_call": round(ms / steps, 4),
            "max_rel": float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))}
    m, nr = 4096, 512
    tri = np.tril(np.random.default_rng(seed + 11).uniform(-1.0, 1.0, (m, m)))
    tri[np.arange(m), np.arange(m)] = 4096.0 + np.random.default_rng(seed + 12).uniform(0.0, 1.0, m)
    rhs = np.random.default_rng(seed + 13).uniform(-1.0, 1.0, (m, nr))
    T, X = DeviceMatrix.from_host(np.asfortranarray(tri)), DeviceMatrix.from_host(np.asfortranarray(rhs))
    blas.dtrsm("L", "L", "N", "N", m, nr, 1.0, T, X, stream=stream)
    stream.synchronize()
    x = X.to_host()
    resid = float(np.max(np.abs(tri @ x - rhs)) / (np.max(np.abs(tri)) * np.max(np.abs(x)) * m))
    ms, _ = _time(lambda: blas.dtrsm("L", "L", "N", "N", m, nr, 1.0, T, X, stream=stream), stream, steps, warmup)
    out["dtrsm_llnn_4096x512"] = {"tflops": round(float(m) * m * nr * steps / (ms / 1e3) / 1e12, 2),
                                  "ms_per_call": round(ms / steps, 4), "residual": resid,
                                  "flops": "M^2 N"}
    out["dgetrf_4096"] = getrf_leg(4096, seed)
    return out


def getrf_leg(n, seed, calls=5):
    """LU with partial pivoting of an n x n uniform[-1,1) matrix (dgetrf
    contract) on device pointers.  Each call factors a fresh copy of the matrix
    (the copy is outside the HIP-event interval); work (2/3) n^3 flop; the
    backward error ||P A - L U||_F / ||A||_F is checked on the last call."""
    import ctypes
    from gemm_avx512f_amd import _lib, blas
    from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix, Event, Stream, fill_uniform

    stream = Stream()
    A0, A = DeviceMatrix(n, n, np.float64), DeviceMatrix(n, n, np.float64)
    fill_uniform(A0, seed + 21, 0, stream=stream)
    stream.synchronize()
    piv, info = DeviceBuffer(4 * n), DeviceBuffer(4)
    nbytes = 8 * n * n
    total = 0.0
    for i in range(calls + 1):
        _lib.check(_lib.lib().gemm_amd_memcpy(ctypes.c_void_p(A.ptr), ctypes.c_void_p(A0.ptr), nbytes, 3), "copy")
        e0, e1 = Event(), Event()
        e0.record(stream)
        blas.dgetrf(n, n, A, piv, info, stream=stream)
        e1.record(stream)
        ms = e0.elapsed_ms(e1)
        if i:   # the first call warms up
            total += ms
    a0, lu = A0.to_host(), A.to_host()
    ipiv, inf = np.zeros(n, np.int32), np.zeros(1, np.int32)
    piv.download(ipiv)
    info.download(inf)
    L, U = np.tril(lu, -1) + np.eye(n), np.triu(lu)
    pa = a0.copy()
    for j, r in enumerate(ipiv):
        if r - 1 != j:
            pa[[j, r - 1]] = pa[[r - 1, j]]
    berr = float(np.linalg.norm(L @ U - pa) / np.linalg.norm(a0))
    ms = total / calls
    return {"tflops": round(2.0 / 3.0 * n ** 3 / (ms / 1e3) / 1e12, 2), "ms_per_call": round(ms, 3),
            "backward_error": berr, "info": int(inf[0]), "flops": "(2/3) n^3"}


def c4_shapes(seed, count=512):
    """M, N, K log-uniform over [32, 2048] rounded to multiples of 8; transposes drawn per problem."""
    rng = np.random.default_rng(seed)
    dims = np.exp(rng.uniform(np.log(32), np.log(2048), size=(count, 3)))
    dims = np.clip(np.rint(dims / 8) * 8, 32, 2048).astype(int)
    ts = rng.integers(0, 2, size=(count, 2))
    return [(int(m), int(n), int(k), "NT"[a], "NT"[b]) for (m, n, k), (a, b) in zip(dims, ts)]


def grouped_leg(replays=20):
    """Batching workloads: 512-problem grouped GEMM (FP32 and bf16 -> FP32) replayed
    from a captured hipGraph, and a batch of 128 complex-double 3M products of
    size 256..1024 (one batched call)."""
    from gemm_avx512f_amd import _lib
    from gemm_avx512f_amd.device import DeviceMatrix, Event, Stream, fill_uniform
    from gemm_avx512f_amd.grouped import GROUPED_BF16, GROUPED_F32, GroupedGemm, zgemm3m_batched

    _lib.check(_lib.lib().gemm_amd_init(), "init")
    out = {}
    stream = Stream()
    for name, dtype in (("fp32", GROUPED_F32), ("bf16_fp32acc", GROUPED_BF16)):
        probs, flops = [], 0.0
        for i, (m, n, k, ta, tb) in enumerate(c4_shapes(2024)):
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            if dtype == GROUPED_F32:
                a, b = DeviceMatrix(ar, ac, np.float32), DeviceMatrix(br, bc, np.float32)
                d = fill_uniform(a, seed=i)
                fill_uniform(b, seed=i, first_draw=d)
            else:
                # bf16 operands: bit patterns of uniform[-1,1) values truncated to bf16, built on the host
                rng = np.random.default_rng(i)
                a = DeviceMatrix.from_host(np.asfortranarray(
                    (rng.uniform(-1, 1, (ar, ac)).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)))
                b = DeviceMatrix.from_host(np.asfortranarray(
                    (rng.uniform(-1, 1, (br, bc)).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)))
            c = DeviceMatrix(m, n, np.float32)
            probs.append(dict(transa=ta, transb=tb, m=m, n=n, k=k, alpha=1.0, beta=0.0, a=a, b=b, c=c))
            flops += 2.0 * m * n * k
        # load balance of the persistent launch under both schedules: per
        # workgroup start/end times of one (uncaptured) launch
        balance = {}
        for sched, dyn in (("static", False), ("dynamic", True)):
            gs = GroupedGemm(dtype, probs, dynamic=dyn)
            gs.timeline(stream)   # warm-up
            st, en, _ = gs.timeline(stream)
            span = float(en.max())
            balance[sched] = {"makespan_us": round(span, 1),
                              "busy_pct": round(100.0 * float((en - st).sum()) / (len(st) * span), 1),
                              "tail_us": round(span - float(np.median(en)), 1)}
            gs.destroy()
        g = GroupedGemm(dtype, probs)   # the library's default schedule (dynamic)
        g.capture(stream)
        for _ in range(3):
            g.replay(stream)
        stream.synchronize()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(replays):
            g.replay(stream)
        e1.record(stream)
        ms = e0.elapsed_ms(e1) / replays
        out[name] = {"problems": len(probs), "gflop": round(flops / 1e9, 2), "ms_per_replay": round(ms, 4),
                     "tflops": round(flops / (ms / 1e3) / 1e12, 2), "tiles": g.tiles, "workgroups": g.grid,
                     "schedule": "dynamic", "load_balance": balance,
                     "load_balance_note": "busy_pct = sum over workgroups of (end - start) / (workgroups x "
                                          "makespan); tail_us = makespan - median workgroup end"}
        g.destroy()
        del probs
    rng = np.random.default_rng(128)
    zp, zflops = [], 0.0
    for i in range(128):
        m, n, k = (int(x) for x in rng.integers(256, 1025, size=3))
        fa, fb = DeviceMatrix(2 * m, k, np.float64), DeviceMatrix(2 * k, n, np.float64)
        d = fill_uniform(fa, seed=9000 + i)
        fill_uniform(fb, seed=9000 + i, first_draw=d)
        za, zb = (DeviceMatrix.__new__(DeviceMatrix) for _ in range(2))
        for z, src, r, cl in ((za, fa, m, k), (zb, fb, k, n)):
            z.rows, z.cols, z.dtype, z.ld, z.buf = r, cl, np.dtype(np.complex128), r, src.buf
        zp.append(dict(transa="N", transb="N", m=m, n=n, k=k, alpha=1.0, beta=0.0, a=za, b=zb,
                       c=DeviceMatrix(m, n, np.complex128)))
        zflops += 8.0 * m * n * k
    zgemm3m_batched(zp, stream)          # warm-up (workspace growth)
    stream.synchronize()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        zgemm3m_batched(zp, stream)
    ms = (time.perf_counter() - t0) / reps * 1e3
    out["zgemm3m_batch"] = {"problems": 128, "ms_per_call": round(ms, 3),
                            "tflops_complex_equiv": round(zflops / (ms / 1e3) / 1e12, 2),
                            "tflops_real_executed": round(0.75 * zflops / (ms / 1e3) / 1e12, 2),
                            "note": "host-timed, each call synchronises; the warm-up call builds the plan, the timed calls (same arguments) reuse it"}
    return out


def physical_cores():
    """Physical cores this process may run on: distinct (package, core) pairs
    of /proc/cpuinfo among the CPUs in the affinity mask, capped by the cgroup
    CPU quota (cpu.max) when one is set."""
    allowed = os.sched_getaffinity(0)
    cores, cpu, pkg = set(), None, "0"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "processor":
                    cpu = int(val)
                elif key == "physical id":
                    pkg = val
                elif key == "core id" and cpu in allowed:
                    cores.add((pkg, val))
    except OSError:
        pass
    n = len(cores) or len(allowed)
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_leg_child(size, reps):
    """Body of the CPU leg (runs in a child process whose OpenMP environment
    was set before the runtime loaded): the reference's dgemm_
    (oracle/_ref/DGEMM.so, -O2 -mavx512f -fopenmp) on this host, or the
    blocked C port of its algorithm where the host lacks AVX-512F."""
    from oracle import oracle as O

    threads = int(os.environ["OMP_NUM_THREADS"])
    kind = "reference" if O.ref_lib("DGEMM.so") is not None else "port"
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
    return {"value": round(2.0 * m * n * k / best / 1e12, 4), "unit": "TFLOP/s", "cores": threads, "kind": kind,
            "sample": "DGEMM %d^3 NN alpha=1 beta=0, seed-0 splitmix64 inputs, best of %d calls after 1 warm-up "
                      "(%.2f s per call), one OpenMP thread per physical core, OMP_PROC_BIND=close "
                      "OMP_PLACES=cores" % (size, reps, best)}


def cpu_leg(size, reps=3):
    """The reference's CPU dgemm_ timed on all physical cores of this host
    (counted by physical_cores), in a child process so that OMP_NUM_THREADS,
    OMP_PROC_BIND=close and OMP_PLACES=cores are in place before the OpenMP
    runtime loads."""
    import subprocess

    cores = physical_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg-child", str(size), str(reps)],
                         env=env, capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError("CPU leg failed: " + out.stderr[-2000:])
    res = json.loads(out.stdout.strip().splitlines()[-1])
    res["host_logical_cpus"] = os.cpu_count()
    return res


def dropin_leg(size, a, b, reps=3):
    """The host-pointer drop-in dgemm_ (the reference's own entry point,
    src/gemm_driver.c:116) at the headline size: copies of A and B into HBM,
    the same kernel, the copy of C back; wall time per call, best of `reps`
    after one warm-up call."""
    from gemm_avx512f_amd import blas

    m = n = k = size
    c = np.zeros((m, n), order="F")
    best = None
    for i in range(reps + 1):
        t0 = time.perf_counter()
        blas.dgemm_("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
        dt = time.perf_counter() - t0
        if i > 0:
            best = dt if best is None else min(best, dt)
    gib = 3.0 * 8 * size * size / 2 ** 30
    return {"ms_per_call": round(best * 1e3, 2), "tflops_incl_pcie": round(2.0 * m * n * k / best / 1e12, 3),
            "host_bytes_moved_gib": round(gib, 3),
            "note": "dgemm_ on host (pageable) arrays: H2D of A and B, kernel, D2H of C (beta = 0 skips the H2D "
                    "of C); wall clock, best of %d after 1 warm-up" % reps}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-emulated", action="store_true", help="skip the emulated (int8) DGEMM legs")
    ap.add_argument("--nslices", type=int, default=0,
                    help="slices per operand of the slice-form emulated DGEMM (0: the library's tuned count for K)")
    ap.add_argument("--nmoduli", type=int, default=0,
                    help="moduli of the modular-form emulated DGEMM, timed at this count and one fewer "
                         "(0: the library's tuned count for K)")
    ap.add_argument("--cpu-size", type=int, default=None, help="CPU leg problem size (default: --size)")
    ap.add_argument("--no-grouped", action="store_true", help="skip the grouped-GEMM / zgemm3m batch leg")
    ap.add_argument("--no-sgemm", action="store_true", help="skip the SGEMM / bf16x split leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the SGEMM 1024^3 / DGEMM 4096^3 / DTRSM 4096x512 leg")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the host-pointer dgemm_ leg (PCIe copies included)")
    ap.add_argument("--cpu-leg-child", nargs=2, type=int, metavar=("SIZE", "REPS"), help=argparse.SUPPRESS)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB first (a known-size kernel that "
                         "calibrates the HBM byte counters when the run is profiled)")
    args = ap.parse_args()
    if args.cpu_leg_child:
        print(json.dumps(cpu_leg_child(*args.cpu_leg_child)), flush=True)
        return
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
    nm_default, ns_default = 0, 0
    if not args.no_emulated:
        from gemm_avx512f_amd.tuning import emulation_counts
        nm_default, ns_default = emulation_counts(size)   # what a call with 0 runs at this K
    nmoduli = args.nmoduli or nm_default
    nslices = args.nslices or ns_default
    variants = () if args.no_emulated else (("crt", nmoduli), ("crt", nmoduli - 1), ("slices", nslices))
    _progress("native and emulated DGEMM %d^3" % size)
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
                       "sample": "%d entries vs long-double dot products" % SAMPLES},
        "wall_s_timed_region": round(wall, 3),
    }
    from gemm_avx512f_amd import _lib
    import ctypes
    tile, splitk = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().gemm_amd_pick_config(1, b"N", b"N", size, size, size, 0, 2, ctypes.byref(tile),
                                               ctypes.byref(splitk)), "pick_config")
    result["config"]["kernel_config"] = {"tile": ["128x128", "64x64", "128x64", "256x64"][tile.value],
                                         "splitk": splitk.value,
                                         "kernels_per_step": 1 if splitk.value == 1 else 2}
    if not args.no_emulated:
        _progress("accuracy against dynamic range")
    accuracy = accuracy_leg(size) if not args.no_emulated else None
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
        })
        # FP64-level: max error no larger than native FP64's on the headline
        # inputs and on every input range of the accuracy leg
        at = {"uniform": bool(e["err"][0] <= legs["native"]["err"][0])}
        for rk in ("r=0", "r=8", "r=32"):
            row = accuracy[rk]
            at[rk] = bool(row[e["accuracy_key"]]["max_rel"] <= row["native_fp64"]["max_rel"])
        variants[-1]["fp64_level_at"] = at
        variants[-1]["fp64_level"] = all(at.values())
        variants[-1]["faster_than_native"] = bool(em_tf > tflops)
        variants[-1]["library_default"] = bool(e["count"] == (nm_default if e["kind"] == "crt" else ns_default))
    if variants:
        # headline emulated figure: the modular form at the library's own count
        # for this K (what a call with nmoduli = 0 runs) when it is at FP64
        # level on every input range; else the fastest variant that is
        ok = [v for v in variants if v["fp64_level"]]
        default = [v for v, e in zip(variants, legs["emulated"]) if e["kind"] == "crt" and v["library_default"]]
        if default and default[0]["fp64_level"]:
            result["emulated"] = default[0]
        else:
            result["emulated"] = max(ok, key=lambda v: v["tflops_fp64_equiv"]) if ok else None
        result["emulated_variants"] = variants
    if accuracy is not None:
        result["accuracy_vs_range"] = accuracy
    if not args.no_dropin:
        _progress("host-pointer dgemm_")
        a, b = legs["host_ab"]
        result["dropin_host_pointer"] = dropin_leg(size, a, b)
    del legs["host_ab"]
    if not args.no_sgemm:
        _progress("SGEMM and bf16x")
        result["sgemm"] = sgemm_leg(size, args.steps, args.warmup)
    if not args.no_grouped:
        _progress("grouped and zgemm3m batch")
        result["batching"] = grouped_leg()
    if not args.no_configs:
        _progress("SGEMM 1024, DGEMM 4096, DTRSM, LU")
        result["configs"] = configs_leg()
    if not args.no_cpu:
        _progress("reference CPU dgemm_")
        try:
            result["cpu_baseline"] = cpu_leg(args.cpu_size or size)
        except Exception as exc:  # reported, never silently replaced
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
