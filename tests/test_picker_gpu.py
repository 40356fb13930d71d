"""Autotuner / learned picker on the GPU.

* the HIP forward, gradient and Adam kernels against the double-precision
  oracle (FP32 tolerances written in each test);
* training is bitwise reproducible and lowers the loss;
* the sweep measures every valid candidate and the heuristic's choice;
* the library's own choice follows table > picker > heuristic, and the HIP
  inference picks the same candidates as the oracle's forward pass;
* live timing: on held-out shapes where the sweep says the heuristic is far
  off, the picker's configuration runs faster than the heuristic's.
"""
import ctypes
import os

import numpy as np
import pytest

from gemm_avx512f_amd import _lib, tuning
from gemm_avx512f_amd.device import DeviceBuffer, DeviceMatrix, fill_uniform, synchronize
from gemm_avx512f_amd import blas
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = O.MLP_NPARAMS


def dev(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = DeviceBuffer(a.nbytes)
    b.upload(a)
    return b


def host(buf, n):
    return buf.download(np.empty(n, dtype=np.float32))


def case(seed, batch):
    rng = np.random.default_rng(seed)
    p = tuning.init_params(seed).astype(np.float64)
    p[O.MLP_NPARAMS - 20:] = rng.normal(0, 0.1, 20)
    x = rng.uniform(-1, 1, (batch, O.MLP_IN))
    t = rng.uniform(0, 3, (batch, 20))
    mask = (rng.uniform(size=(batch, 20)) < 0.8).astype(np.float64)
    return p, x, t, mask


def test_forward_matches_oracle(gpu_lib):
    p, x, _, _ = case(1, 37)
    y = DeviceBuffer(37 * 20 * 4)
    dp, dx = dev(p), dev(x)
    _lib.check(gpu_lib.gemm_amd_mlp_forward(dp.ptr, dx.ptr, 37, y.ptr, None), "forward")
    synchronize()
    got = host(y, 37 * 20).reshape(37, 20)
    want = O.mlp_forward(p.astype(np.float32).astype(np.float64), x.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.max(np.abs(want)))


@pytest.mark.parametrize("batch", [1, 64, 300])
def test_grad_matches_oracle(gpu_lib, batch):
    p, x, t, mask = case(2 + batch, batch)
    p32, x32, t32 = (a.astype(np.float32) for a in (p, x, t))
    dp, dx, dt, dm = dev(p32), dev(x32), dev(t32), dev(mask)
    g, loss = DeviceBuffer(P * 4), DeviceBuffer(4)
    _lib.check(gpu_lib.gemm_amd_mlp_grad(dp.ptr, dx.ptr, dt.ptr, dm.ptr, batch, g.ptr, loss.ptr, None), "grad")
    synchronize()
    gg = host(g, P)
    want, wl = O.mlp_grad(p32.astype(np.float64), x32.astype(np.float64), t32.astype(np.float64), mask)
    # FP32 sums of <= 300 terms: error well inside 1e-4 of the gradient's scale
    assert np.max(np.abs(gg - want)) <= 1e-4 * np.max(np.abs(want)), np.max(np.abs(gg - want))
    assert host(loss, 1)[0] == pytest.approx(wl, rel=1e-5)


def test_adam_matches_oracle(gpu_lib):
    rng = np.random.default_rng(5)
    p = rng.normal(size=P).astype(np.float32)
    dp, dm, dv = dev(p), dev(np.zeros(P)), dev(np.zeros(P))
    pr, mr, vr = p.astype(np.float64), np.zeros(P), np.zeros(P)
    for step in (1, 2, 3):
        g = rng.normal(size=P).astype(np.float32)
        dg = dev(g)
        _lib.check(gpu_lib.gemm_amd_mlp_adam(dp.ptr, dg.ptr, dm.ptr, dv.ptr, P, 1e-3, 0.9, 0.999, 1e-8, step, None),
                   "adam")
        O.adam(pr, g.astype(np.float64), mr, vr, 1e-3, 0.9, 0.999, 1e-8, step)
    synchronize()
    assert np.max(np.abs(host(dp, P) - pr)) <= 1e-6


def test_training_reproducible_and_learns(gpu_lib):
    # a learnable synthetic target: the best candidate is a function of (m, k)
    s = tuning.sample_shapes(11, 2048)
    x = tuning.features(s)
    t = np.abs(np.subtract.outer(x[:, 0] + x[:, 2], np.linspace(-2, 2, 20))).astype(np.float32)
    mask = np.ones_like(t)
    p1, l1 = tuning.train(x, t, mask, epochs=40, batch=128, lr=3e-3, seed=3)
    p2, l2 = tuning.train(x, t, mask, epochs=40, batch=128, lr=3e-3, seed=3)
    assert np.array_equal(p1, p2) and np.array_equal(l1, l2)
    assert l1[-1] < 0.1 * l1[0], (l1[0], l1[-1])
    y = O.mlp_forward(p1.astype(np.float64), x.astype(np.float64))
    pick = np.argmin(y, axis=1)
    # the picked candidate's target is close to the best one (neighbours differ by 4/19)
    assert np.mean(t[np.arange(len(t)), pick] - t.min(axis=1)) < 0.1


def test_sweep_measures_valid_candidates(gpu_lib):
    ms = tuning.sweep(tuning.OP_DGEMM, False, True, 300, 200, 100, pad=5)
    valid = tuning.valid_mask(1, 300, 200, 100)
    assert np.all(ms[:20][valid] > 0) and np.all(ms[:20][~valid] == -1) and ms[20] > 0
    ms = tuning.sweep(tuning.OP_CGEMM3M, True, False, 64, 96, 2048)
    assert np.all(ms[:20][tuning.valid_mask(2, 64, 96, 2048)] > 0) and ms[20] > 0


def test_library_choice_order(gpu_lib):
    params = tuning.init_params(9)
    tuning.picker_load(params)
    tuning.picker_set_mode(1)
    shp = np.zeros(1, dtype=tuning.SHAPE_DTYPE)
    shp[0] = (0, 0, 1, 777, 555, 333, 3)
    want = int(tuning.picker_choose(shp)[0])
    t, s = tuning.pick_config(0, False, True, 777, 555, 333, 3, mode=2)
    assert tuning.candidate_index(t, s) == want
    # a table entry overrides the picker for exactly that shape
    ms = np.full((1, 21), -1.0, dtype=np.float32)
    other = (want + 5) % 20 if tuning.valid_mask(0, 777, 555, 333)[(want + 5) % 20] else 0
    ms[0, other] = 1.0
    tuning.table_set(shp, ms)
    assert tuning.candidate_index(*tuning.pick_config(0, False, True, 777, 555, 333, 3, mode=2)) == other
    tuning.table_clear()
    tuning.picker_set_mode(0)
    assert tuning.pick_config(0, False, True, 777, 555, 333, 3, mode=2) == tuning.pick_config(
        0, False, True, 777, 555, 333, 3, mode=0)
    # GEMM results do not depend on the mode (only the schedule changes)
    a, b = DeviceMatrix(777, 333, np.float32), DeviceMatrix(555, 333, np.float32)
    d = fill_uniform(a, seed=1)
    fill_uniform(b, seed=1, first_draw=d)
    c0, c1 = DeviceMatrix(777, 555, np.float32), DeviceMatrix(777, 555, np.float32)
    blas.sgemm("N", "T", 777, 555, 333, 1.0, a, b, 0.0, c0)
    tuning.picker_set_mode(1)
    blas.sgemm("N", "T", 777, 555, 333, 1.0, a, b, 0.0, c1)
    synchronize()
    x0, x1 = c0.to_host(), c1.to_host()
    assert np.max(np.abs(x0 - x1)) <= 1e-5 * np.max(np.abs(x0))
    _restore_default_picker()


def _restore_default_picker():
    if os.path.exists(os.path.join(tuning.DATA, "picker.npz")):
        tuning.picker_load(np.load(os.path.join(tuning.DATA, "picker.npz"))["params"])
        tuning.picker_set_mode(1)
    else:
        tuning.picker_set_mode(0)


HELD = os.path.join(tuning.DATA, "sweep_heldout.npz")
WEIGHTS = os.path.join(tuning.DATA, "picker.npz")
needs_data = pytest.mark.skipif(not (os.path.exists(HELD) and os.path.exists(WEIGHTS)),
                                reason="sweep data / trained weights not generated yet")


ALIGNED = os.path.join(tuning.DATA, "sweep_aligned_heldout.npz")


@needs_data
@pytest.mark.parametrize("data", [HELD, ALIGNED], ids=["heldout", "aligned"])
def test_hip_inference_matches_oracle_picks(gpu_lib, data):
    d = np.load(data, allow_pickle=False)
    shapes = d["shapes"]
    params = np.load(WEIGHTS)["params"]
    tuning.picker_load(params)
    got = tuning.picker_choose(shapes)
    y = O.mlp_forward(params.astype(np.float64), tuning.features(shapes).astype(np.float64))
    valid = tuning.valid_mask(shapes["op"], shapes["m"], shapes["n"], shapes["k"])
    yv = np.where(valid, y, np.inf)
    want = tuning.choose(y, valid)
    # disagreements only where a comparison is within FP32 noise: the two best
    # predictions, or the split-K tie-break margin against its threshold
    srt = np.sort(yv, axis=1)
    rows = np.arange(len(yv))
    best = np.argmin(yv, axis=1)
    one = (best // 5) * 5
    margin = np.where(valid[rows, one], yv[rows, one] - yv[rows, best], np.inf)
    close = ((srt[:, 1] - srt[:, 0]) < 1e-4) | (np.abs(margin - tuning.PICK_TIE) < 1e-4)
    assert np.all((got == want) | close)
    assert np.mean(got == want) > 0.99


@needs_data
def test_picker_beats_heuristic_live(gpu_lib):
    """Shapes where the sweep shows the heuristic >= 1.5x off the picker's choice: time both again."""
    d = np.load(HELD, allow_pickle=False)
    shapes, ms = d["shapes"], d["ms"]
    tuning.picker_load(np.load(WEIGHTS)["params"])
    choice = tuning.picker_choose(shapes)
    t_pick = ms[np.arange(len(ms)), choice]
    gain = ms[:, 20] / t_pick
    idx = [i for i in np.argsort(-gain) if ms[i, 20] < 20.0][:8]    # bounded run time
    assert len(idx) == 8 and gain[idx[-1]] >= 1.5
    wins = 0
    for i in idx:
        s = shapes[i]
        again = tuning.sweep(int(s["op"]), bool(s["ta"]), bool(s["tb"]), int(s["m"]), int(s["n"]), int(s["k"]),
                             int(s["pad"]))
        wins += again[choice[i]] < again[20]
    assert wins >= 7, wins
