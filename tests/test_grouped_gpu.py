"""Grouped GEMM (one persistent launch over many problems) and the batched
zgemm3m on the GPU.

* every problem of a group matches the long-double oracle within the GEMM
  bound (FP32, FP64; bf16 operands are exact in FP32 products, so their bound
  is the FP32 one);
* FP32/FP64 problems are bitwise equal to the single-problem kernel run with
  the same tile shape and slab depth and no split-K (same k order, same
  epilogue);
* a plan captured into a hipGraph and replayed several times, with eager work
  and synchronisations between the replays, reproduces the eager launch bit
  for bit every time;
* the 512-problem workload (M, N, K log-uniform over [32, 2048], multiples of
  8, random transposes) matches per-problem single GEMMs;
* the zgemm3m batch matches the 4M complex oracle (small) and the one-problem
  zgemm3m (128 problems of 256..1024).
"""
import numpy as np
import pytest

from gemm_avx512f_amd import blas, grouped
from gemm_avx512f_amd.device import DeviceMatrix, Stream, fill_uniform, synchronize
from gemm_avx512f_amd.grouped import GROUPED_BF16, GROUPED_F32, GROUPED_F64, GroupedGemm, zgemm3m_batched
from oracle import oracle as O
from tests.util import opmat

pytestmark = pytest.mark.gpu

# tile classes per dtype: small 64x64 tiles, large tiles for problems whose M
# and N (FP64: N half of it) reach the plan's threshold, by default bf16 96,
# FP64 128 and none for FP32
SMALL_TILE = (64, 64)
LARGE_TILE = {GROUPED_F32: (128, 128), GROUPED_BF16: (128, 128), GROUPED_F64: (128, 64)}
DEFAULT_LARGE_MIN = {GROUPED_F32: 1 << 30, GROUPED_BF16: 96, GROUPED_F64: 128}


def expected_tiles(dtype, problems, large_min=None):
    t = DEFAULT_LARGE_MIN[dtype] if not large_min else large_min
    total = 0
    for p in problems:
        large = p["m"] >= t and p["n"] >= (t // 2 if dtype == GROUPED_F64 else t)
        bm, bn = LARGE_TILE[dtype] if large else SMALL_TILE
        rem = p["m"] % bm
        if dtype == GROUPED_F64 and large and p["m"] > bm and 0 < rem <= SMALL_TILE[0]:
            # FP64: the rows of whole large tiles, then the rest on small tiles
            total += (p["m"] // bm) * -(-p["n"] // bn) + -(-rem // SMALL_TILE[0]) * -(-p["n"] // SMALL_TILE[1])
            continue
        total += -(-p["m"] // bm) * -(-p["n"] // bn)
    return total


SAME_SINGLE_TILE = {GROUPED_F32: 1, GROUPED_F64: 1}   # TILE_64x64


def f32_to_bf16_bits(x):
    """Truncate FP32 values to bf16 bit patterns (the values stay in [-1, 1))."""
    return (np.ascontiguousarray(x, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16).reshape(x.shape)


def bf16_bits_to_f64(h):
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def random_problems(dtype, count, seed, lo, hi, mult=1, pad_max=3, zero_dims=False):
    """Host problems with random shapes, transposes, padding and scalars."""
    rng = np.random.default_rng(seed)
    out = []
    for p in range(count):
        m, n, k = (int(round(np.exp(rng.uniform(np.log(lo), np.log(hi))) / mult)) * mult for _ in range(3))
        m, n, k = max(m, mult), max(n, mult), max(k, mult)
        if zero_dims and p % 7 == 3:
            k = 0
        if zero_dims and p % 11 == 5:
            m = 0
        ta, tb = rng.choice(["N", "T"]), rng.choice(["N", "T"])
        pad = int(rng.integers(0, pad_max + 1))
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        s = O.Stream(seed * 1000 + p)
        if dtype == GROUPED_F64:
            a, b, c = s.f64(ar, ac, ar + pad), s.f64(br, bc, br + pad), s.f64(m, n, m + pad)
        else:
            a, b, c = s.f32(ar, ac, ar + pad), s.f32(br, bc, br + pad), s.f32(m, n, m + pad)
            if dtype == GROUPED_BF16:
                a, b = np.asfortranarray(f32_to_bf16_bits(a)), np.asfortranarray(f32_to_bf16_bits(b))
        alpha = float(rng.choice([1.0, -0.75, 1.5]))
        beta = float(rng.choice([0.0, 1.0, 0.5]))
        out.append(dict(transa=ta, transb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, ha=a, hb=b, hc=c,
                        rows_a=ar, rows_b=br))
    return out


def upload(problems):
    for p in problems:
        p["a"] = DeviceMatrix.from_padded(p["ha"], p["rows_a"])
        p["b"] = DeviceMatrix.from_padded(p["hb"], p["rows_b"])
        p["c"] = DeviceMatrix.from_padded(p["hc"], p["m"])
    return problems


def oracle_check(dtype, p, out):
    m, n, k = p["m"], p["n"], p["k"]
    c = p["hc"]
    assert np.array_equal(out[m:, :], c[m:, :]), "rows beyond M were written"
    if m == 0 or n == 0:
        return
    if dtype == GROUPED_BF16:
        a, b = bf16_bits_to_f64(p["ha"]), bf16_bits_to_f64(p["hb"])
    else:
        a, b = p["ha"].astype(np.float64), p["hb"].astype(np.float64)
    a, b = np.asfortranarray(a), np.asfortranarray(b)
    ref = np.asfortranarray(c.astype(np.float64))
    O.dgemm(p["transa"], p["transb"], m, n, k, p["alpha"], a, a.shape[0], b, b.shape[0], p["beta"], ref,
            ref.shape[0])
    A, B = opmat(a, p["transa"], m, k), opmat(b, p["transb"], k, n)
    scale = abs(p["alpha"]) * (np.abs(A) @ np.abs(B)) + abs(p["beta"]) * np.abs(c[:m, :n]) + 1e-300
    u = 2.0 ** -53 if dtype == GROUPED_F64 else 2.0 ** -24
    err = np.max(np.abs(out[:m, :n] - ref[:m, :n]) / scale)
    assert err <= 4 * (k + 2) * u, (p["transa"], p["transb"], m, n, k, err)


# "ragged": arbitrary sizes and padded leading dimensions (range-checked element
# loads); "chunked": sizes and leading dimensions multiples of 8 (whole 16-byte
# chunk loads through buffer resources, two slabs ahead), still ragged against
# the 64 x 64 tiles and the slab depth
SHAPES = {"ragged": dict(mult=1, pad_max=3), "chunked": dict(mult=8, pad_max=0)}


@pytest.mark.parametrize("shapes", sorted(SHAPES))
@pytest.mark.parametrize("dtype", [GROUPED_F32, GROUPED_BF16, GROUPED_F64], ids=["f32", "bf16", "f64"])
def test_grouped_vs_oracle(dtype, shapes):
    P = upload(random_problems(dtype, 40, seed=11 + dtype, lo=1, hi=300, zero_dims=True, **SHAPES[shapes]))
    g = GroupedGemm(dtype, P)
    assert g.tiles == expected_tiles(dtype, P)
    g.run()
    synchronize()
    for p in P:
        oracle_check(dtype, p, p["c"].to_host_padded())
    g.destroy()


@pytest.mark.parametrize("shapes", sorted(SHAPES))
@pytest.mark.parametrize("dtype", [GROUPED_F32, GROUPED_F64], ids=["f32", "f64"])
def test_grouped_bitwise_equals_single_gemm(dtype, shapes):
    P = upload(random_problems(dtype, 30, seed=21 + dtype, lo=8, hi=500, **SHAPES[shapes]))
    g = GroupedGemm(dtype, P)
    g.run()
    single = blas.sgemm if dtype == GROUPED_F32 else blas.dgemm
    for p in P:
        c2 = DeviceMatrix.from_padded(p["hc"], p["m"])
        single(p["transa"], p["transb"], p["m"], p["n"], p["k"], p["alpha"], p["a"], p["b"], p["beta"], c2,
               tile=SAME_SINGLE_TILE[dtype], splitk=1)
        p["c_single"] = c2
    synchronize()
    for p in P:
        assert np.array_equal(p["c"].to_host_padded(), p["c_single"].to_host_padded()), \
            (p["transa"], p["transb"], p["m"], p["n"], p["k"])


@pytest.mark.parametrize("dtype", [GROUPED_F32, GROUPED_BF16, GROUPED_F64], ids=["f32", "bf16", "f64"])
def test_tile_classes_bitwise_equal(dtype):
    # every problem on large tiles where it qualifies, against every problem on
    # small tiles: same k order and epilogue, so the same bits
    P = upload(random_problems(dtype, 40, seed=61 + dtype, lo=16, hi=700))
    results = {}
    for t in (32, 1 << 30):
        for p in P:
            p["c"].buf.upload(p["hc"])
        g = GroupedGemm(dtype, P, large_min=t)
        assert g.tiles == expected_tiles(dtype, P, large_min=t)
        g.run()
        synchronize()
        results[t] = [p["c"].to_host_padded() for p in P]
        g.destroy()
    assert any(p["m"] >= 32 and p["n"] >= 32 for p in P)
    for p, big, small in zip(P, results[32], results[1 << 30]):
        assert np.array_equal(big, small), (p["transa"], p["transb"], p["m"], p["n"], p["k"])
    for p, got in zip(P, results[32]):
        oracle_check(dtype, p, got)


@pytest.mark.parametrize("dtype", [GROUPED_F32, GROUPED_BF16], ids=["f32", "bf16"])
def test_graph_replay_bitwise_stable(dtype):
    P = upload(random_problems(dtype, 24, seed=31 + dtype, lo=16, hi=400))
    for p in P:
        p["beta"] = 0.5          # replays must read C afresh each time
    g = GroupedGemm(dtype, P)
    g.run()
    synchronize()
    eager = [p["c"].to_host_padded() for p in P]
    s = Stream()
    g.capture(s)
    # unrelated eager work between replays, on the default stream
    x = DeviceMatrix.from_host(np.asfortranarray(np.ones((96, 96), np.float32)))
    y = DeviceMatrix(96, 96, np.float32)
    for rep in range(4):
        for p in P:
            p["c"].buf.upload(p["hc"])
        blas.sgemm("N", "N", 96, 96, 96, 1.0, x, x, 0.0, y)
        synchronize()
        g.replay(s)
        s.synchronize()
        for p, want in zip(P, eager):
            assert np.array_equal(p["c"].to_host_padded(), want), ("replay", rep, p["m"], p["n"], p["k"])
        assert np.all(y.to_host() == 96.0)
    g.destroy()


@pytest.mark.parametrize("dtype", [GROUPED_F32, GROUPED_BF16, GROUPED_F64], ids=["f32", "bf16", "f64"])
def test_dynamic_schedule_bitwise_equal_and_replays(dtype):
    # tiles claimed from an atomic counter are the same tiles computed the same
    # way: the same bits as the static stride; the counter's reset kernel sits
    # in the captured graph, so every replay starts again from tile 0
    P = upload(random_problems(dtype, 30, seed=71 + dtype, lo=16, hi=600))
    for p in P:
        p["beta"] = 0.5
    g = GroupedGemm(dtype, P, dynamic=False)
    g.run()
    synchronize()
    want = [p["c"].to_host_padded() for p in P]
    g.destroy()
    gd = GroupedGemm(dtype, P, dynamic=True)
    for p in P:
        p["c"].buf.upload(p["hc"])
    gd.run()
    synchronize()
    for p, w in zip(P, want):
        assert np.array_equal(p["c"].to_host_padded(), w), ("dynamic", p["m"], p["n"], p["k"])
    s = Stream()
    gd.capture(s)
    for rep in range(3):
        for p in P:
            p["c"].buf.upload(p["hc"])
        synchronize()
        gd.replay(s)
        s.synchronize()
        for p, w in zip(P, want):
            assert np.array_equal(p["c"].to_host_padded(), w), ("replay", rep, p["m"], p["n"], p["k"])
    gd.destroy()


@pytest.mark.parametrize("dynamic", [False, True], ids=["static", "dynamic"])
def test_timeline_accounts_for_every_tile(dynamic):
    P = upload(random_problems(GROUPED_F32, 40, seed=81, lo=16, hi=500))
    g = GroupedGemm(GROUPED_F32, P, dynamic=dynamic)
    start, end, tiles = g.timeline()
    assert len(start) == g.grid and g.grid >= 1
    assert int(tiles.sum()) == g.tiles
    assert np.all(end >= start) and start.min() == 0 and end.max() < 1e6   # microseconds
    if not dynamic:   # workgroup w takes tiles w, w + grid, ...
        assert list(tiles) == [len(range(w, g.tiles, g.grid)) for w in range(g.grid)]
    for p in P:   # the timed launch computed the products
        oracle_check(GROUPED_F32, p, p["c"].to_host_padded())
    g.destroy()


def test_replay_without_capture_raises():
    P = upload(random_problems(GROUPED_F32, 2, seed=41, lo=8, hi=32))
    g = GroupedGemm(GROUPED_F32, P)
    with pytest.raises(grouped._lib.GemmAmdError):
        g.replay()
    g.destroy()


def test_grouped_illegal_argument_returns_null():
    P = upload(random_problems(GROUPED_F32, 3, seed=42, lo=8, hi=32))
    P[1]["c"].ld = 0                   # ldc < max(1, m): parameter 13 of problem 1
    with pytest.raises(grouped._lib.GemmAmdError, match="problem 1"):
        GroupedGemm(GROUPED_F32, P)


def c4_shapes(seed, count=512):
    """M, N, K log-uniform over [32, 2048] rounded to multiples of 8, transposes random."""
    rng = np.random.default_rng(seed)
    dims = np.exp(rng.uniform(np.log(32), np.log(2048), size=(count, 3)))
    dims = np.clip(np.rint(dims / 8) * 8, 32, 2048).astype(int)
    ts = rng.integers(0, 2, size=(count, 2))
    return [(int(m), int(n), int(k), "NT"[a], "NT"[b]) for (m, n, k), (a, b) in zip(dims, ts)]


def test_c4_512_problems_fp32_match_single_gemm_and_graph():
    P = []
    for i, (m, n, k, ta, tb) in enumerate(c4_shapes(1234)):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        a, b, c = DeviceMatrix(ar, ac, np.float32), DeviceMatrix(br, bc, np.float32), DeviceMatrix(m, n, np.float32)
        d = fill_uniform(a, seed=5000 + i)
        d = fill_uniform(b, seed=5000 + i, first_draw=d)
        P.append(dict(transa=ta, transb=tb, m=m, n=n, k=k, alpha=1.0, beta=0.0, a=a, b=b, c=c))
    g = GroupedGemm(GROUPED_F32, P)
    s = Stream()
    g.capture(s)
    g.replay(s)
    s.synchronize()
    bad = []
    for i, p in enumerate(P):
        c2 = DeviceMatrix(p["m"], p["n"], np.float32)
        blas.sgemm(p["transa"], p["transb"], p["m"], p["n"], p["k"], 1.0, p["a"], p["b"], 0.0, c2, tile=SAME_SINGLE_TILE[GROUPED_F32],
                   splitk=1)
        if i % 16 == 0 or i < 4:
            if not np.array_equal(p["c"].to_host(), c2.to_host()):
                bad.append(i)
    assert not bad, bad
    g.destroy()


def test_c4_512_problems_bf16_vs_sampled_oracle():
    """bf16 operands (bit patterns of uniform[-1,1) truncated to bf16): every 32nd problem is checked
    elementwise on a sample of entries against exact long-double dot products."""
    P = []
    rng = np.random.default_rng(77)
    for i, (m, n, k, ta, tb) in enumerate(c4_shapes(4321)):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        ha = np.asfortranarray(f32_to_bf16_bits(rng.uniform(-1, 1, size=(ar, ac)).astype(np.float32)))
        hb = np.asfortranarray(f32_to_bf16_bits(rng.uniform(-1, 1, size=(br, bc)).astype(np.float32)))
        P.append(dict(transa=ta, transb=tb, m=m, n=n, k=k, alpha=1.0, beta=0.0, a=DeviceMatrix.from_host(ha),
                      b=DeviceMatrix.from_host(hb), c=DeviceMatrix(m, n, np.float32), ha=ha, hb=hb))
    g = GroupedGemm(GROUPED_BF16, P)
    s = Stream()
    g.capture(s)
    g.replay(s)
    s.synchronize()
    for i in range(0, len(P), 32):
        p = P[i]
        m, n, k = p["m"], p["n"], p["k"]
        A = opmat(bf16_bits_to_f64(p["ha"]), p["transa"], m, k)
        B = opmat(bf16_bits_to_f64(p["hb"]), p["transb"], k, n)
        out = p["c"].to_host()
        rows, cols = rng.integers(0, m, 64), rng.integers(0, n, 64)
        for r, c in zip(rows, cols):
            exact = np.dot(A[r, :].astype(np.longdouble), B[:, c].astype(np.longdouble))
            scale = np.dot(np.abs(A[r, :]), np.abs(B[:, c]))
            assert abs(float(out[r, c] - exact)) <= 4 * (k + 2) * 2.0 ** -24 * scale, (i, r, c)
    g.destroy()


def zcase(count, seed, lo, hi, ops):
    rng = np.random.default_rng(seed)
    P = []
    for p in range(count):
        m, n, k = (int(x) for x in rng.integers(lo, hi + 1, size=3))
        ta, tb = ops[p % len(ops)]
        s = O.Stream(seed * 100 + p)
        a = s.c128(m, k, m + 1) if ta == "N" else s.c128(k, m, k + 1)
        b = s.c128(k, n, k) if tb == "N" else s.c128(n, k, n)
        c = s.c128(m, n, m + 2)
        alpha = complex(rng.uniform(-1, 1), rng.uniform(-1, 1))
        beta = complex(rng.uniform(-1, 1), rng.uniform(-1, 1)) if p % 3 else 0j
        P.append(dict(transa=ta, transb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, ha=a, hb=b, hc=c,
                      a=DeviceMatrix.from_padded(a, a.shape[0]), b=DeviceMatrix.from_padded(b, b.shape[0]),
                      c=DeviceMatrix.from_padded(c, m)))
    return P


def op(x, t):
    return x if t == "N" else (x.T if t == "T" else x.conj().T)


def test_zgemm3m_batched_vs_oracle():
    P = zcase(12, 9, 1, 150, [("N", "N"), ("T", "N"), ("N", "C"), ("C", "T"), ("C", "C")])
    zgemm3m_batched(P)
    synchronize()
    for p in P:
        m, n, k = p["m"], p["n"], p["k"]
        a, b, c = p["ha"], p["hb"], p["hc"]
        ref = c.copy(order="F")
        O.zgemm(p["transa"], p["transb"], m, n, k, p["alpha"], a, a.shape[0], b, b.shape[0], p["beta"], ref,
                c.shape[0])
        out = p["c"].to_host_padded()
        assert np.array_equal(out[m:, :], c[m:, :])
        A = op(a, p["transa"])[:m, :k]
        B = op(b, p["transb"])[:k, :n]
        bound = abs(p["alpha"]) * ((np.abs(A.real) + np.abs(A.imag)) @ (np.abs(B.real) + np.abs(B.imag))) \
            + abs(p["beta"]) * np.abs(c[:m, :n]) + 1e-300
        err = np.max(np.abs(out[:m, :n] - ref[:m, :n]) / bound)
        assert err <= 8 * (k + 4) * 2.0 ** -53, (p["transa"], p["transb"], m, n, k, err)


def test_zgemm3m_batched_repeat_and_changed_scalar():
    """A repeated batch (the library reuses its plan) gives the same bits; a
    changed alpha is not served by the old plan (alpha doubled doubles C)."""
    P = zcase(7, 21, 1, 90, [("N", "N"), ("T", "C"), ("C", "N")])
    for p in P:
        p["beta"] = 0j
    zgemm3m_batched(P)
    synchronize()
    first = [p["c"].to_host_padded().copy() for p in P]
    zgemm3m_batched(P)
    synchronize()
    for p, f in zip(P, first):
        assert np.array_equal(p["c"].to_host_padded(), f)
    P[0]["alpha"] = 2 * P[0]["alpha"]
    zgemm3m_batched(P)
    synchronize()
    m = P[0]["m"]
    assert np.array_equal(P[0]["c"].to_host_padded()[:m], 2 * first[0][:m])
    for p, f in zip(P[1:], first[1:]):
        assert np.array_equal(p["c"].to_host_padded(), f)


def test_zgemm3m_batch_128_matches_single():
    """128 complex-double problems of size 256..1024 against one zgemm3m call per problem."""
    rng = np.random.default_rng(128)
    P, S = [], []
    for i in range(128):
        m, n, k = (int(x) for x in rng.integers(256, 1025, size=3))
        dims = dict(transa="N", transb="N", m=m, n=n, k=k, alpha=1.0 + 0.5j, beta=0j)
        # fill the interleaved storage as real data: real and imag parts uniform[-1,1)
        fa = DeviceMatrix(2 * m, k, np.float64)
        fb = DeviceMatrix(2 * k, n, np.float64)
        d = fill_uniform(fa, seed=9000 + i)
        fill_uniform(fb, seed=9000 + i, first_draw=d)
        za, zb = _as_complex(fa, m, k), _as_complex(fb, k, n)
        P.append(dict(dims, a=za, b=zb, c=DeviceMatrix(m, n, np.complex128)))
        S.append(dict(dims, a=za, b=zb, c=DeviceMatrix(m, n, np.complex128)))
    zgemm3m_batched(P)
    for q in S:
        blas.zgemm3m("N", "N", q["m"], q["n"], q["k"], q["alpha"], q["a"], q["b"], q["beta"], q["c"])
    synchronize()
    for i in range(0, 128, 8):
        x, y = P[i]["c"].to_host(), S[i]["c"].to_host()
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        assert rel <= 4 * (P[i]["k"] + 2) * 2.0 ** -53, (i, rel)


def _as_complex(real_mat, rows, cols):
    """View a (2*rows x cols) FP64 DeviceMatrix as a rows x cols complex128 DeviceMatrix (same storage)."""
    z = DeviceMatrix.__new__(DeviceMatrix)
    z.rows, z.cols, z.dtype, z.ld = rows, cols, np.dtype(np.complex128), rows
    z.buf = real_mat.buf
    return z
