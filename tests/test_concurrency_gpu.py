"""Device-pointer calls from several host threads, each on its own stream, run
at the same time without sharing workspace: split-K DGEMMs (whose partial
sums live in library workspace) of growing sizes, so that workspace grows
while the other thread's kernels are queued, and DTRSM solves (diagonal
inverses and trailing-update partials in workspace).  Every result is checked
against the oracle after all threads have finished.  The reference's dgemm_
allocates its scratch per call and is reentrant (src/gemm_driver.c:133-140,
189-193); this is the same guarantee for the stream-ordered entry points."""
import threading

import numpy as np
import pytest

from gemm_avx512f_amd import blas
from gemm_avx512f_amd.device import DeviceMatrix, Stream
from oracle import oracle as O
from tests.util import gemm_bound, gemm_case, opmat

pytestmark = pytest.mark.gpu

# per thread: (m, n, k) of its split-K DGEMMs, growing so that the split-K
# partials need a larger buffer on almost every call
SHAPES = {
    0: [(96, 80, 512), (160, 144, 640), (256, 200, 768), (320, 288, 896), (416, 384, 1024)],
    1: [(128, 64, 448), (192, 176, 576), (272, 240, 832), (352, 336, 960), (448, 400, 1088)],
}
TRSM = {0: [(320, 96), (520, 160)], 1: [(384, 128), (600, 200)]}


def _gemm_jobs(tid):
    jobs = []
    for i, (m, n, k) in enumerate(SHAPES[tid]):
        a, b, c, lda, ldb, ldc = gemm_case(np.float64, 7000 + 100 * tid + i, "N", "T", m, n, k, pad=1)
        ref = c.copy(order="F")
        O.dgemm("N", "T", m, n, k, 1.5, a, lda, b, ldb, 0.5, ref, ldc)
        absab = 1.5 * (np.abs(opmat(a, "N", m, k)) @ np.abs(opmat(b, "T", k, n)))
        denom = absab + 0.5 * np.abs(c[:m, :n]) + np.finfo(np.float64).tiny
        jobs.append(dict(m=m, n=n, k=k, dA=DeviceMatrix.from_padded(a, a.shape[0]),
                         dB=DeviceMatrix.from_padded(b, b.shape[0]), dC=DeviceMatrix.from_padded(c, m),
                         ref=ref[:m, :n], denom=denom))
    return jobs


def _trsm_jobs(tid):
    jobs = []
    for i, (m, n) in enumerate(TRSM[tid]):
        s = O.Stream(8000 + 100 * tid + i)
        a = np.tril(s.f64(m, m))
        a[np.arange(m), np.arange(m)] = m + np.abs(a[np.arange(m), np.arange(m)])
        b = s.f64(m, n)
        jobs.append(dict(m=m, n=n, a=a, b=b, dA=DeviceMatrix.from_host(np.asfortranarray(a)),
                         dX=DeviceMatrix.from_host(np.asfortranarray(b))))
    return jobs


def test_concurrent_streams_do_not_share_workspace(gpu_lib):
    gemm_jobs = {t: _gemm_jobs(t) for t in SHAPES}
    trsm_jobs = {t: _trsm_jobs(t) for t in TRSM}
    start = threading.Barrier(len(SHAPES))
    errors = []

    def worker(tid):
        try:
            stream = Stream()
            start.wait()
            # everything is enqueued back to back; nothing waits until the end
            for i, j in enumerate(gemm_jobs[tid]):
                blas.dgemm("N", "T", j["m"], j["n"], j["k"], 1.5, j["dA"], j["dB"], 0.5, j["dC"], stream=stream,
                           tile=1, splitk=4)
                if i < len(trsm_jobs[tid]):
                    t = trsm_jobs[tid][i]
                    blas.dtrsm("L", "L", "N", "N", t["m"], t["n"], 1.0, t["dA"], t["dX"], stream=stream)
            stream.synchronize()
        except Exception as exc:  # reported by the main thread
            errors.append((tid, repr(exc)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in SHAPES]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in threads), "a worker thread did not finish"
    assert not errors, errors
    for tid, jobs in gemm_jobs.items():
        for j in jobs:
            out = j["dC"].to_host_padded()[:j["m"], :j["n"]]
            err = float(np.max(np.abs(out - j["ref"]) / j["denom"]))
            assert err <= gemm_bound(np.float64, j["k"]), (tid, j["m"], j["n"], j["k"], err)
    for tid, jobs in trsm_jobs.items():
        for t in jobs:
            x = t["dX"].to_host()
            m = t["m"]
            resid = np.max(np.abs(t["a"] @ x - t["b"])) / (np.max(np.abs(t["a"])) * np.max(np.abs(x)) * m)
            assert resid < 1e-15, (tid, m, t["n"], resid)


def test_same_stream_from_two_threads(gpu_lib):
    # two threads sharing one stream: their calls are ordered by the stream and
    # each call's workspace is held until its kernels are enqueued
    stream = Stream()
    jobs = {t: _gemm_jobs(t) for t in SHAPES}
    errors = []

    def worker(tid):
        try:
            for j in jobs[tid]:
                blas.dgemm("N", "T", j["m"], j["n"], j["k"], 1.5, j["dA"], j["dB"], 0.5, j["dC"], stream=stream,
                           tile=1, splitk=4)
        except Exception as exc:
            errors.append((tid, repr(exc)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in SHAPES]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    stream.synchronize()
    assert not errors, errors
    for tid, js in jobs.items():
        for j in js:
            out = j["dC"].to_host_padded()[:j["m"], :j["n"]]
            err = float(np.max(np.abs(out - j["ref"]) / j["denom"]))
            assert err <= gemm_bound(np.float64, j["k"]), (tid, j["m"], j["n"], j["k"], err)
