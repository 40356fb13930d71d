"""Autotuner and learned configuration picker (libgemm_amd.so).

* :func:`sweep` times all 20 candidate configurations (tile shape x split-K)
  of one problem on the GPU; the sweep is exhaustive, so it is both the
  fallback for shapes it has seen (:func:`table_set`) and the judge of the
  picker's regret.
* The picker is a 34-64-64-20 ReLU perceptron predicting, per candidate,
  log2(time / best time).  It is trained on the GPU by the library's own HIP
  kernels (:func:`train`), and run by its HIP inference kernel
  (:func:`picker_choose`, and inside every GEMM call for unseen shapes).
* :func:`tune_emulation` picks the moduli / slice counts of the emulated
  DGEMM per K (the counts the library uses when called with 0), and
  :func:`tune_grouped` the tile-class threshold of a grouped workload.

Command line (runs on a GPU host)::

    python -m gemm_avx512f_amd.tuning sweep --seed 1 --start 0 --count 1000 --out sweep.npz
    python -m gemm_avx512f_amd.tuning train --data a.npz b.npz --out picker.npz
    python -m gemm_avx512f_amd.tuning tune-emulation --out counts.json --write-inc
"""
import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import _lib

OP_SGEMM, OP_DGEMM, OP_CGEMM3M = 0, 1, 2
OP_NAMES = {OP_SGEMM: "sgemm", OP_DGEMM: "dgemm", OP_CGEMM3M: "cgemm3m"}
NCFG = 20
MLP_IN, MLP_HID, MLP_OUT = 34, 64, 20
MLP_NPARAMS = MLP_IN * MLP_HID + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID * MLP_OUT + MLP_OUT
TARGET_CLIP = 3.0          # targets are log2(t / t_best) clipped at 8x slower
HERE = os.path.dirname(os.path.abspath(__file__))
DATA = os.path.join(HERE, "data")
WEIGHTS_INC = os.path.join(HERE, "csrc", "picker_weights.inc")

SHAPE_DTYPE = np.dtype([("op", np.int32), ("ta", np.int32), ("tb", np.int32), ("m", np.int32), ("n", np.int32),
                        ("k", np.int32), ("pad", np.int32)])


def candidates():
    t = (ctypes.c_int * NCFG)()
    s = (ctypes.c_int * NCFG)()
    n = _lib.lib().gemm_amd_tune_candidates(t, s)
    return [(t[i], s[i]) for i in range(n)]


def candidate_index(tile, splitk):
    return tile * 5 + int(np.log2(splitk))


PICK_TIE = 0.02   # log2 units: see choose()


def choose(y, valid):
    """Candidate chosen from predicted costs y (count x NCFG), the rule the
    library applies after inference: the valid minimum, except that a split-K
    candidate gives way to the same tile without split-K when that one is
    predicted within PICK_TIE (no reduce pass, no workspace; the network's
    resolution is a few percent)."""
    yy = np.where(valid, y, np.inf)
    best = np.argmin(yy, axis=1)
    rows = np.arange(len(best))
    one = (best // 5) * 5
    take = valid[rows, one] & (yy[rows, one] - yy[rows, best] <= PICK_TIE)
    return np.where(take, one, best)


def valid_mask(op, m, n, k):
    """Candidates the library considers valid (each K partition >= 64 deep, partials <= 1 GiB)."""
    m, n, k = (np.asarray(v, dtype=np.float64) for v in (m, n, k))
    eb = np.where(np.asarray(op) == OP_DGEMM, 8.0, 4.0)
    out = np.zeros(m.shape + (NCFG,), dtype=bool)
    for i in range(NCFG):
        s = 1 << (i % 5)
        out[..., i] = True if s == 1 else (k >= 64 * s) & (m * n * s * eb <= 2.0 ** 30)
    return out


# ------------------------------------------------------------------ shapes
def sample_shapes(seed, count):
    """Seeded GEMM problems: op balanced over SGEMM/DGEMM/CGEMM3M; M, N, K
    log-uniform over [16, 16384]; 20% with skinny K in [16, 128]; 20% with
    M or N in [16, 64]; random transposes; lda padding 0 (half) or 1..63."""
    rng = np.random.default_rng(seed)

    def lu(lo, hi, size):
        return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size))).astype(np.int32)

    s = np.zeros(count, dtype=SHAPE_DTYPE)
    s["op"] = np.arange(count) % 3
    s["m"], s["n"], s["k"] = lu(16, 16384, count), lu(16, 16384, count), lu(16, 16384, count)
    kind = rng.uniform(size=count)
    skinny = kind < 0.2
    s["k"][skinny] = lu(16, 128, int(skinny.sum()))
    small = (kind >= 0.2) & (kind < 0.4)
    which = rng.integers(0, 2, count)
    sm_m, sm_n = small & (which == 0), small & (which == 1)
    s["m"][sm_m] = lu(16, 64, int(sm_m.sum()))
    s["n"][sm_n] = lu(16, 64, int(sm_n.sum()))
    s["ta"], s["tb"] = rng.integers(0, 2, count), rng.integers(0, 2, count)
    pad = rng.integers(1, 64, count)
    s["pad"] = np.where(rng.uniform(size=count) < 0.5, 0, pad)
    return s


def sample_shapes_aligned(seed, count):
    """Seeded GEMM problems of the sizes programs and benchmarks favour: each
    of M, N, K a power of two in [64, 16384] (half the draws) or a multiple
    of 64, log-uniform over the same range; op balanced; random transposes;
    lda padding 0 in 3 of 4 problems (the 16-byte load path), else 1..63.
    Here the tile grid often fills whole waves of workgroups exactly, which
    is what the wave-fill features resolve."""
    rng = np.random.default_rng(seed)
    s = np.zeros(count, dtype=SHAPE_DTYPE)
    s["op"] = np.arange(count) % 3
    for key in ("m", "n", "k"):
        pow2 = 2 ** rng.integers(6, 15, count)
        mult = np.maximum(64, np.rint(np.exp(rng.uniform(np.log(64), np.log(16384), count)) / 64) * 64)
        s[key] = np.where(rng.uniform(size=count) < 0.5, pow2, mult).astype(np.int32)
    s["ta"], s["tb"] = rng.integers(0, 2, count), rng.integers(0, 2, count)
    pad = rng.integers(1, 64, count)
    s["pad"] = np.where(rng.uniform(size=count) < 0.75, 0, pad)
    return s


CUS = 256          # MI355X compute units (the library passes the device's own count)
# macro-tile (BM, BN) of the four tile shapes, FP32 and FP64 (gemm_mfma.hip kTilesF32 / kTilesF64)
TILE_BMBN = {4: ((128, 128), (64, 64), (128, 64), (256, 128)), 8: ((128, 128), (64, 64), (128, 64), (256, 64))}


def k_partitions(k, splitk):
    """K partitions a split-K request runs as (each a whole number of 32-deep slabs)."""
    k = np.asarray(k, dtype=np.int64)
    chunk = (k + splitk - 1) // splitk
    chunk = (chunk + 31) // 32 * 32
    return np.where(k > 0, (k + chunk - 1) // np.maximum(chunk, 1), 1)


def vec_path(shapes, ldb_pad=0):
    """1 where the kernels take the 16-byte vector-load path (library's gemm_vec_path;
    sweeps store B with ldb = rows of op(B))."""
    s = shapes
    op, m, n, k = (s[f].astype(np.int64) for f in ("op", "m", "n", "k"))
    lda = np.where(s["ta"] != 0, k, m) + s["pad"]
    ldb = np.where(s["tb"] != 0, n, k) + ldb_pad
    vw = np.where(op == OP_DGEMM, 2, 4)
    real = (lda % vw == 0) & (ldb % vw == 0)
    planes = (m % 4 == 0) & (k % 4 == 0)        # CGEMM3M products read packed planes (ld M, K)
    return np.where(op == OP_CGEMM3M, planes, real)


def features(shapes, cus=CUS):
    """The network's inputs (same formula as the library's mlp_features)."""
    s = shapes
    f = np.zeros((len(s), MLP_IN), dtype=np.float64)
    f[:, 0] = (np.log2(np.maximum(s["m"], 1)) - 9.0) / 5.0
    f[:, 1] = (np.log2(np.maximum(s["n"], 1)) - 9.0) / 5.0
    f[:, 2] = (np.log2(np.maximum(s["k"], 1)) - 9.0) / 5.0
    f[:, 3] = s["ta"]
    f[:, 4] = s["tb"]
    for o in range(3):
        f[:, 5 + o] = s["op"] == o
    f[:, 8] = (s["pad"] % 64) / 64.0
    f[:, 9] = vec_path(s)
    m = np.maximum(s["m"].astype(np.int64), 1)
    n = np.maximum(s["n"].astype(np.int64), 1)
    f64 = s["op"] == OP_DGEMM
    for t in range(4):
        bm = np.where(f64, TILE_BMBN[8][t][0], TILE_BMBN[4][t][0])
        bn = np.where(f64, TILE_BMBN[8][t][1], TILE_BMBN[4][t][1])
        tm, tn = (m + bm - 1) // bm, (n + bn - 1) // bn
        for j in range(5):
            w = tm * tn * k_partitions(s["k"], 1 << j)
            f[:, 10 + 5 * t + j] = w / (((w + cus - 1) // cus) * cus)
        f[:, 30 + t] = (m * n) / (tm * bm * tn * bn)
    return f.astype(np.float32)


def targets(ms):
    """log2(t_c / t_best) clipped at TARGET_CLIP, and the mask of measured candidates."""
    ms = np.asarray(ms, dtype=np.float64)[:, :NCFG]
    ok = ms > 0
    best = np.min(np.where(ok, ms, np.inf), axis=1, keepdims=True)
    t = np.where(ok, np.minimum(np.log2(np.where(ok, ms, 1.0) / best), TARGET_CLIP), 0.0)
    return t.astype(np.float32), ok.astype(np.float32)


def regret(ms, choice):
    """t(choice) / t(best) - 1 per shape (choice = candidate index, or NCFG for the heuristic slot)."""
    ms = np.asarray(ms, dtype=np.float64)
    best = np.min(np.where(ms[:, :NCFG] > 0, ms[:, :NCFG], np.inf), axis=1)
    picked = ms[np.arange(len(ms)), choice]
    return picked / best - 1.0


# ------------------------------------------------------------------ GPU sweep
def sweep(op, ta, tb, m, n, k, pad=0):
    ms = (ctypes.c_float * (NCFG + 1))()
    rc = _lib.lib().gemm_amd_tune_sweep(op, b"T" if ta else b"N", b"T" if tb else b"N", m, n, k, pad, ms)
    _lib.check(rc, "tune_sweep")
    return np.array(ms[:], dtype=np.float32)


def sweep_shapes(shapes, log=None, every=100, ops=None):
    """Sweep every shape (or only those whose op is in `ops`; the others are left at -2)."""
    out = np.full((len(shapes), NCFG + 1), -2.0, dtype=np.float32)
    t0 = time.time()
    for i, s in enumerate(shapes):
        if ops is not None and int(s["op"]) not in ops:
            continue
        out[i] = sweep(int(s["op"]), bool(s["ta"]), bool(s["tb"]), int(s["m"]), int(s["n"]), int(s["k"]),
                       int(s["pad"]))
        if log is not None and (i + 1) % every == 0:
            log("swept %d/%d shapes in %.1f s" % (i + 1, len(shapes), time.time() - t0))
    return out


# ------------------------------------------------------------------ picker
def _shape_args(shapes):
    s = shapes
    cnt = len(s)
    ints = lambda a: (ctypes.c_int * max(cnt, 1))(*[int(v) for v in a])  # noqa: E731
    ta = ctypes.create_string_buffer(b"".join(b"T" if v else b"N" for v in s["ta"]) or b"N")
    tb = ctypes.create_string_buffer(b"".join(b"T" if v else b"N" for v in s["tb"]) or b"N")
    return [ints(s["op"]), ta, tb, ints(s["m"]), ints(s["n"]), ints(s["k"]), ints(s["pad"])]


def library_features(shapes, cus=CUS):
    """The library's own feature rows for these shapes (host computation, no GPU)."""
    out = np.zeros((len(shapes), MLP_IN), dtype=np.float32)
    args = _shape_args(shapes)
    _lib.check(_lib.lib().gemm_amd_picker_features(len(shapes), *args, cus, out.ctypes.data_as(ctypes.c_void_p)),
               "picker_features")
    return out


def picker_load(params):
    p = np.ascontiguousarray(params, dtype=np.float32)
    _lib.check(_lib.lib().gemm_amd_picker_load(p.ctypes.data_as(ctypes.c_void_p), p.size), "picker_load")


def picker_set_mode(mode):
    _lib.check(_lib.lib().gemm_amd_picker_set_mode(mode), "picker_set_mode")


def picker_choose(shapes):
    """Candidate index chosen by the picker (HIP inference) for each shape."""
    args = _shape_args(shapes)
    out = (ctypes.c_int * max(len(shapes), 1))()
    _lib.check(_lib.lib().gemm_amd_picker_choose(len(shapes), *args, out), "picker_choose")
    return np.array(out[: len(shapes)], dtype=np.int64)


def pick_config(op, ta, tb, m, n, k, pad=0, mode=2):
    t, s = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().gemm_amd_pick_config(op, b"T" if ta else b"N", b"T" if tb else b"N", m, n, k, pad, mode,
                                               ctypes.byref(t), ctypes.byref(s)), "pick_config")
    return t.value, s.value


def table_set(shapes, ms):
    """Install the sweep's winners for these exact shapes (they override the picker)."""
    ms = np.asarray(ms)
    choice = np.argmin(np.where(ms[:, :NCFG] > 0, ms[:, :NCFG], np.inf), axis=1)
    args = _shape_args(shapes)
    ch = (ctypes.c_int * max(len(shapes), 1))(*[int(c) for c in choice])
    _lib.check(_lib.lib().gemm_amd_tune_table_set(len(shapes), *args, ch), "tune_table_set")


def table_clear():
    _lib.lib().gemm_amd_tune_table_clear()


# ------------------------------------------------------------------ training (GPU kernels)
def init_params(seed):
    """He-uniform weights, zero biases (float32, the library's flat layout)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((MLP_IN, MLP_HID), (MLP_HID, MLP_HID), (MLP_HID, MLP_OUT)):
        lim = np.sqrt(6.0 / fan_in)
        parts += [rng.uniform(-lim, lim, size=fan_in * fan_out), np.zeros(fan_out)]
    return np.concatenate(parts).astype(np.float32)


SIZE_WEIGHT = 2.0   # see size_weights


def size_weights(shapes, alpha=SIZE_WEIGHT):
    """Per-shape loss weights 1 + alpha * max(0, log2(M N K) - 30): a shape's
    share of the loss grows with its work, so the network resolves the large
    shapes (where a wrong pick costs the most time) more finely.  alpha = 2
    measured best on the held-out sweep: mean regret 0.97% (0.98% unweighted),
    p90 2.96% (3.34%), total-time regret 1.94% (2.34%)."""
    lf = np.log2(np.maximum(shapes["m"].astype(np.float64) * shapes["n"] * shapes["k"], 1.0))
    return 1.0 + alpha * np.maximum(0.0, lf - 30.0)


def train(x, t, mask, epochs=300, batch=256, lr=2e-3, seed=0, params=None):
    """Train on the GPU with the library's HIP kernels; returns (params, per-epoch loss)."""
    from .device import DeviceBuffer
    x, t, mask = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, t, mask))
    p0 = init_params(seed) if params is None else np.ascontiguousarray(params, dtype=np.float32)
    bufs = {}
    for name, arr in (("p", p0), ("m", np.zeros_like(p0)), ("v", np.zeros_like(p0)), ("x", x), ("t", t),
                      ("mask", mask)):
        bufs[name] = DeviceBuffer(arr.nbytes)
        bufs[name].upload(arr)
    losses = np.zeros(epochs, dtype=np.float32)
    rc = _lib.lib().gemm_amd_mlp_train(bufs["p"].ptr, bufs["m"].ptr, bufs["v"].ptr, bufs["x"].ptr, bufs["t"].ptr,
                                       bufs["mask"].ptr, len(x), batch, epochs, lr, seed,
                                       losses.ctypes.data_as(ctypes.c_void_p))
    _lib.check(rc, "mlp_train")
    out = bufs["p"].download(np.empty_like(p0))
    return out, losses


def write_weights_inc(params, path=WEIGHTS_INC, note=""):
    """Write the trained parameters as the C include compiled into the library."""
    p = np.asarray(params, dtype=np.float32)
    assert p.size == MLP_NPARAMS
    lines = ["// Trained weights of the configuration picker (flat layout of mlp.hpp),",
             "// written by gemm_avx512f_amd/tuning.py." + (" " + note if note else ""),
             "static const float kPickerWeights[%d] = {" % p.size]
    vals = ["%.8ef" % float(v) if np.isfinite(v) else "0.0f" for v in p]   # 9 digits round-trip FP32
    for i in range(0, len(vals), 8):
        lines.append("    " + ", ".join(vals[i:i + 8]) + ",")
    lines += ["};", "static const int kPickerWeightsCount = %d;" % p.size, ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


# ------------------------------------------------------------------ emulation counts
COUNTS_INC = os.path.join(HERE, "csrc", "emulation_counts.inc")
CRT_COUNTS = (16, 15, 14, 13, 12, 11, 10, 9, 8)
SLICE_COUNTS = (8, 7, 6, 5, 4, 3)


def _max_rel(c, ref):
    return float(np.max(np.abs(c - ref)) / np.max(np.abs(ref)))


def tune_emulation(ks=tuple(2 ** e for e in range(6, 18)), ranges=(0, 8, 32), size=512, seed=31, log=None):
    """Per K, the smallest CRT moduli count and slice count whose max relative
    error stays at or below the native FP64 kernel's for A, B entries
    u * 2^e, e uniform in [-r, r], at every r.  Errors are taken against the
    16-moduli CRT result, whose own error is far below both (about 2^-55
    relative at K = 8192 against native's 2^-49), so the comparison needs no
    host reference; every product runs on the GPU through the library.
    Returns {k: {"moduli": n, "slices": s, "errors": {...}}}."""
    from . import blas
    from .device import FILL_F64_RANGE, DeviceMatrix, fill_uniform, synchronize

    out = {}
    for k in ks:
        A, B = DeviceMatrix(size, k, np.float64), DeviceMatrix(k, size, np.float64)
        C = DeviceMatrix(size, size, np.float64)
        ok_crt = {n: True for n in CRT_COUNTS}
        ok_sl = {n: True for n in SLICE_COUNTS}
        errs = {}
        for r in ranges:
            d = fill_uniform(A, seed + r, 0, kind=FILL_F64_RANGE, range_r=r)
            fill_uniform(B, seed + r, d, kind=FILL_F64_RANGE, range_r=r)
            blas.dgemm_ozaki_crt("N", "N", size, size, k, 1.0, A, B, 0.0, C, nmoduli=16)
            synchronize()
            ref = C.to_host()
            blas.dgemm("N", "N", size, size, k, 1.0, A, B, 0.0, C)
            synchronize()
            nat = _max_rel(C.to_host(), ref)
            e = {"native": nat}
            for n in CRT_COUNTS[1:]:
                if _lib.lib().gemm_amd_ozaki_crt_bits(n, k) < 1:
                    ok_crt[n] = False
                    continue
                blas.dgemm_ozaki_crt("N", "N", size, size, k, 1.0, A, B, 0.0, C, nmoduli=n)
                synchronize()
                e["crt%d" % n] = _max_rel(C.to_host(), ref)
                ok_crt[n] = ok_crt[n] and e["crt%d" % n] <= nat
            for n in SLICE_COUNTS:
                blas.dgemm_ozaki("N", "N", size, size, k, 1.0, A, B, 0.0, C, nslices=n)
                synchronize()
                e["slices%d" % n] = _max_rel(C.to_host(), ref)
                ok_sl[n] = ok_sl[n] and e["slices%d" % n] <= nat
            errs["r=%d" % r] = e
        # the smallest count that passes, with every larger count passing too
        nm = 16
        for n in CRT_COUNTS[1:]:
            if not ok_crt[n]:
                break
            nm = n
        ns = 8
        for n in SLICE_COUNTS:
            if not ok_sl[n]:
                break
            ns = n
        if not ok_sl[8]:
            ns = 8   # nothing fewer than the most accurate count
        out[int(k)] = {"moduli": nm, "slices": ns, "errors": errs}
        if log:
            log("K %d: %d moduli, %d slices" % (k, nm, ns))
        del A, B, C
    return out


def counts_tables(tuned):
    """33-entry tables (index ceil(log2 K)) from tune_emulation's result: each
    bucket takes the count of the nearest tuned K at or above it (the counts
    grow with K), the largest tuned K's beyond the range."""
    ks = sorted(tuned)
    mod, sl = [], []
    for b in range(33):
        kk = 1 << b
        sel = [k for k in ks if k >= kk]
        src = tuned[sel[0]] if sel else tuned[ks[-1]]
        if not sel:   # beyond the tuned range: the most accurate counts
            src = {"moduli": 16, "slices": 8}
        mod.append(int(src["moduli"]))
        sl.append(int(src["slices"]))
    return mod, sl


def write_counts_inc(tuned, path=COUNTS_INC, note=""):
    mod, sl = counts_tables(tuned)
    fmt = lambda v: "\n".join("    " + ", ".join(str(x) for x in v[i:i + 16]) + "," for i in range(0, 33, 16))  # noqa: E731
    text = ("// Moduli (CRT form) and slices (slice form) of the emulated DGEMM per\n"
            "// ceil(log2 K), the smallest counts whose max error stayed at or below the\n"
            "// native FP64 kernel's on inputs of dynamic range 2^0, 2^8 and 2^32, written\n"
            "// by gemm_avx512f_amd/tuning.py (tune-emulation)." + (" " + note if note else "") + "\n"
            "static const int kCrtAutoModuli[33] = {\n" + fmt(mod) + "\n};\n"
            "static const int kSliceAutoCount[33] = {\n" + fmt(sl) + "\n};\n")
    with open(path, "w") as f:
        f.write(text)


def emulation_counts(k):
    """The (moduli, slices) the library uses for this K when called with 0."""
    nm, ns = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.lib().gemm_amd_emulation_counts(k, ctypes.byref(nm), ctypes.byref(ns)), "emulation_counts")
    return nm.value, ns.value


def emulation_counts_set(tuned):
    mod, sl = counts_tables(tuned)
    a, b = (ctypes.c_int * 33)(*mod), (ctypes.c_int * 33)(*sl)
    _lib.check(_lib.lib().gemm_amd_emulation_counts_set(a, b, 33), "emulation_counts_set")


# ------------------------------------------------------------------ grouped tile classes
GROUPED_THRESHOLDS = (64, 96, 128, 192, 256, 1 << 30)


def tune_grouped(dtype, problems, thresholds=GROUPED_THRESHOLDS, replays=10, stream=None):
    """Time one grouped workload (GroupedGemm problems) for each tile-class
    threshold (smallest M and N of a problem on the large tiles; 2^30 = small
    tiles only) as a captured graph's replay; returns (best threshold,
    {threshold: ms per replay})."""
    from .device import Event, Stream
    from .grouped import GroupedGemm

    stream = stream or Stream()
    times = {}
    for t in thresholds:
        g = GroupedGemm(dtype, problems, large_min=t)
        try:
            g.capture(stream)
            for _ in range(2):
                g.replay(stream)
            stream.synchronize()
            e0, e1 = Event(), Event()
            e0.record(stream)
            for _ in range(replays):
                g.replay(stream)
            e1.record(stream)
            times[int(t)] = e0.elapsed_ms(e1) / replays
        finally:
            g.destroy()
    return min(times, key=times.get), times


def regret_report(shapes, ms, choice):
    """Median / mean / 90th-percentile regret of `choice` and of the heuristic, overall and per op."""
    r_pick = regret(ms, np.asarray(choice))
    r_heur = regret(ms, np.full(len(ms), NCFG))
    out = {"shapes": int(len(ms))}
    groups = [("all", np.ones(len(ms), dtype=bool))] + [(OP_NAMES[o], shapes["op"] == o) for o in OP_NAMES]
    for name, sel in groups:
        out[name] = {key: {"median": float(np.median(r[sel])), "mean": float(np.mean(r[sel])),
                           "p90": float(np.percentile(r[sel], 90)),
                           "exact_best": float(np.mean(r[sel] <= 0.0))}
                     for key, r in (("picker", r_pick), ("heuristic", r_heur))}
    return out


def merge_sweeps(old_path, new_path, out_path):
    """Replace the rows of an older sweep that a newer partial sweep (--ops) measured again."""
    old, new = np.load(old_path, allow_pickle=False), np.load(new_path, allow_pickle=False)
    if not np.array_equal(old["shapes"], new["shapes"]):
        raise ValueError("the two sweeps cover different shape lists")
    ms = old["ms"].copy()
    fresh = new["ms"][:, 0] != -2.0
    ms[fresh] = new["ms"][fresh]
    np.savez_compressed(out_path, shapes=old["shapes"], ms=ms, seed=old["seed"], start=old["start"])
    return int(fresh.sum())


def load_dataset(paths):
    shapes, ms = [], []
    for p in paths:
        d = np.load(p, allow_pickle=False)
        shapes.append(d["shapes"])
        ms.append(d["ms"])
    return np.concatenate(shapes), np.concatenate(ms)


# ------------------------------------------------------------------ command line
def _main(argv):
    ap = argparse.ArgumentParser(prog="python -m gemm_avx512f_amd.tuning")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--seed", type=int, required=True)
    s.add_argument("--total", type=int, default=None, help="size of the seeded shape list (default start+count)")
    s.add_argument("--start", type=int, default=0)
    s.add_argument("--count", type=int, required=True)
    s.add_argument("--out", required=True)
    s.add_argument("--ops", type=int, nargs="+", default=None, help="sweep only these ops (others stay -2)")
    s.add_argument("--aligned", action="store_true", help="shapes from sample_shapes_aligned")
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--out", required=True)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--size-weight", type=float, default=SIZE_WEIGHT, help="alpha of size_weights (0: unweighted)")
    ev = sub.add_parser("evaluate")
    ev.add_argument("--weights", required=True)
    ev.add_argument("--data", nargs="+", required=True)
    ev.add_argument("--out", default=None)
    te = sub.add_parser("tune-emulation", help="moduli / slice counts of the emulated DGEMM per K")
    te.add_argument("--out", required=True)
    te.add_argument("--write-inc", action="store_true", help="also rewrite csrc/emulation_counts.inc")
    one = sub.add_parser("shape", help="time every candidate of one shape and print TFLOP/s per candidate")
    one.add_argument("--op", type=int, required=True)
    one.add_argument("--ta", type=int, default=0)
    one.add_argument("--tb", type=int, default=0)
    one.add_argument("--m", type=int, required=True)
    one.add_argument("--n", type=int, required=True)
    one.add_argument("--k", type=int, required=True)
    one.add_argument("--pad", type=int, default=0)
    a = ap.parse_args(argv)
    log = lambda msg: print(msg, flush=True)  # noqa: E731
    if a.cmd == "tune-emulation":
        import json
        _lib.check(_lib.lib().gemm_amd_init(), "init")
        tuned = tune_emulation(log=log)
        with open(a.out, "w") as f:
            json.dump({str(k): v for k, v in tuned.items()}, f, indent=1)
        if a.write_inc:
            write_counts_inc(tuned)
        log("wrote " + a.out)
        return 0
    if a.cmd == "shape":
        ms = sweep(a.op, bool(a.ta), bool(a.tb), a.m, a.n, a.k, a.pad)
        fl = 2.0 * a.m * a.n * a.k * (3 if a.op == OP_CGEMM3M else 1)
        cands = candidates()
        best = int(np.argmin(np.where(ms[:NCFG] > 0, ms[:NCFG], np.inf)))
        one = np.zeros(1, dtype=SHAPE_DTYPE)
        one["op"], one["ta"], one["tb"], one["m"], one["n"], one["k"], one["pad"] = (
            a.op, a.ta, a.tb, a.m, a.n, a.k, a.pad)
        learned = int(picker_choose(one)[0])
        cells = ["%s%s tile %d split %d: %.1f" % ("*" if c == best else " ", "P" if c == learned else " ",
                                                  cands[c][0], cands[c][1], fl / (ms[c] * 1e-3) / 1e12)
                 for c in range(NCFG) if ms[c] > 0]
        log("%s ta=%d tb=%d %dx%dx%d pad %d (TFLOP/s; * = fastest, P = learned picker; heuristic's choice %.1f):" % (
            OP_NAMES[a.op], a.ta, a.tb, a.m, a.n, a.k, a.pad, fl / (ms[NCFG] * 1e-3) / 1e12))
        for c in cells:
            log("  " + c)
        return 0
    if a.cmd == "evaluate":
        import json
        shapes, ms = load_dataset(a.data)
        picker_load(np.load(a.weights)["params"])
        stats = regret_report(shapes, ms, picker_choose(shapes))
        log(json.dumps(stats))
        if a.out:
            with open(a.out, "w") as f:
                json.dump(stats, f, indent=1)
        return 0
    if a.cmd == "sweep":
        total = a.total or (a.start + a.count)
        sampler = sample_shapes_aligned if a.aligned else sample_shapes
        shapes = sampler(a.seed, total)[a.start:a.start + a.count]
        log("sweeping %d shapes (seed %d, %d..%d)" % (len(shapes), a.seed, a.start, a.start + len(shapes)))
        ms = sweep_shapes(shapes, log=log, ops=set(a.ops) if a.ops else None)
        np.savez_compressed(a.out, shapes=shapes, ms=ms, seed=a.seed, start=a.start, aligned=int(a.aligned))
        log("wrote " + a.out)
    else:
        shapes, ms = load_dataset(a.data)
        x = features(shapes)
        tt, mask = targets(ms)
        mask = mask * size_weights(shapes, a.size_weight)[:, None]
        t0 = time.time()
        params, losses = train(x, tt, mask, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
        log("trained %d epochs on %d shapes in %.1f s; loss %.4f -> %.4f" % (a.epochs, len(x), time.time() - t0,
                                                                           losses[0], losses[-1]))
        np.savez_compressed(a.out, params=params, losses=losses)
    return 0


if __name__ == "__main__":
    sys.exit(_main(sys.argv[1:]))
