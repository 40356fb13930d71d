"""Autotuner / learned picker, CPU side.

* the perceptron oracle (oracle/mlp_oracle.c) is pinned by hand-computed
  outputs, by central finite differences of its loss, and by Adam's closed
  form on the first step;
* the feature, target and regret helpers and the seeded shape sampler;
* with the committed sweep data and trained weights: the picker's median
  regret on held-out shapes is within 5% of the exhaustive sweep and lower
  than the built-in heuristic's (picks computed with the oracle's forward
  pass from the same weights the library compiles in).
"""
import os
import re

import numpy as np
import pytest

from gemm_avx512f_amd import tuning
from oracle import oracle as O

P = O.MLP_NPARAMS
NIN = O.MLP_IN
OFF_W1, OFF_B1 = 0, NIN * 64
OFF_W2, OFF_B2 = OFF_B1 + 64, OFF_B1 + 64 + 64 * 64
OFF_W3 = OFF_B2 + 64
OFF_B3 = OFF_W3 + 64 * 20


def test_layout_size():
    assert NIN == 34 == tuning.MLP_IN
    assert P == 7700 == tuning.MLP_NPARAMS == OFF_B3 + 20


def test_forward_hand_computed():
    p = np.zeros(P)
    p[OFF_B3:OFF_B3 + 20] = np.arange(20)            # output = b3 when everything else is 0
    x = np.random.default_rng(0).uniform(-1, 1, (3, NIN))
    assert np.array_equal(O.mlp_forward(p, x), np.tile(np.arange(20.0), (3, 1)))
    # identity paths: h1[j] = relu(x[j]) (j < 9), h2 = h1, y[0] = sum_j h2[j], y[1] = -2 * h2[0]
    p = np.zeros(P)
    for j in range(9):
        p[OFF_W1 + j * 64 + j] = 1.0
    for j in range(64):
        p[OFF_W2 + j * 64 + j] = 1.0
    for j in range(9):
        p[OFF_W3 + j * 20 + 0] = 1.0
    p[OFF_W3 + 0 * 20 + 1] = -2.0
    x = np.zeros((1, NIN))
    x[0, [0, 1, 2, 8]] = [0.5, -1.0, 2.0, 0.25]
    y = O.mlp_forward(p, x)
    assert y[0, 0] == 0.5 + 2.0 + 0.25          # the negative input is cut by the ReLU
    assert y[0, 1] == -1.0
    assert np.all(y[0, 2:] == 0)


def _random_case(seed, batch=5):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, P)
    x = rng.uniform(-1, 1, (batch, NIN))
    t = rng.uniform(0, 3, (batch, 20))
    mask = (rng.uniform(size=(batch, 20)) < 0.8).astype(np.float64)
    return p, x, t, mask


def test_grad_matches_finite_differences():
    p, x, t, mask = _random_case(1)
    g, loss = O.mlp_grad(p, x, t, mask)
    assert loss == pytest.approx(O.mlp_loss(p, x, t, mask), rel=1e-14)
    h = 1e-6
    rng = np.random.default_rng(2)
    # every bias and output weight, and a random third of the hidden weights
    idx = np.concatenate([np.arange(OFF_B1, OFF_B1 + 64), np.arange(OFF_B2, OFF_B2 + 64), np.arange(OFF_W3, P),
                          rng.choice(OFF_B1, 200, replace=False), rng.choice(np.arange(OFF_W2, OFF_B2), 600,
                                                                               replace=False)])
    num = np.empty(len(idx))
    for j, i in enumerate(idx):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        num[j] = (O.mlp_loss(pp, x, t, mask) - O.mlp_loss(pm, x, t, mask)) / (2 * h)
    assert np.max(np.abs(num - g[idx])) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    assert np.count_nonzero(g[idx]) > len(idx) // 2     # the check is not vacuous


def test_masked_outputs_carry_no_gradient():
    p, x, t, mask = _random_case(3)
    mask[:] = 0.0
    mask[:, 7] = 1.0
    g, _ = O.mlp_grad(p, x, t, mask)
    w3 = g[OFF_W3:OFF_B3].reshape(64, 20)
    assert np.all(w3[:, np.arange(20) != 7] == 0) and np.all(g[OFF_B3:][np.arange(20) != 7] == 0)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    p0 = rng.normal(size=50)
    g = rng.normal(size=50)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    O.adam(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, 1)
    # bias-corrected moments equal g and g^2 after one step
    assert np.allclose(p, p0 - 1e-2 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_features_formula_points():
    s = np.zeros(4, dtype=tuning.SHAPE_DTYPE)
    s[0] = (1, 1, 0, 512, 16384, 16, 65)
    s[1] = (2, 0, 1, 16, 512, 512, 0)
    s[2] = (0, 0, 0, 2048, 2048, 2048, 0)
    s[3] = (0, 0, 0, 2185, 2185, 2048, 0)
    f = tuning.features(s)
    assert f.shape == (4, 34)
    assert np.allclose(f[0, :9], [0, 1, -1, 1, 0, 0, 1, 0, 1 / 64], atol=1e-7)
    assert np.allclose(f[1, :9], [-1, 0, 0, 0, 1, 0, 0, 1, 0], atol=1e-7)
    # FP64, lda = K + 65 = 81 is odd: scalar loads.  Every tile grid is a
    # whole number of 256-CU waves and K = 16 is one partition at any split.
    assert f[0, 9] == 0 and np.all(f[0, 10:] == 1)
    # CGEMM3M products read packed planes (M = 16, K = 512: vector path);
    # 128x128 / 256x128 tiles: 4 of them, x K partitions 1..16 (K = 512 gives
    # 512 / 32 slabs, so every split is honoured); 64x64 / 128x64: 8 tiles
    assert f[1, 9] == 1
    few = np.array([4, 8, 16, 32, 64]) / 256
    assert np.allclose(f[1, 10:30], np.concatenate([few, 2 * few, 2 * few, few]), atol=1e-7)
    assert np.allclose(f[1, 30:], [16 / 128, 16 / 64, 16 / 128, 16 / 256], atol=1e-7)
    # 2048^2 FP32 fills the chip with 128x128 tiles exactly (256 of them);
    # 2185^2 needs 18 x 18 = 324, a second wave 68/256 full
    assert f[2, 10] == 1 and np.isclose(f[3, 10], 324 / 512)
    assert f[2, 30] == 1 and np.isclose(f[3, 30], 2185 ** 2 / 2304 ** 2)
    assert np.isclose(f[3, 11], 648 / 768)             # split 2: 648 workgroups, three waves


def test_k_partitions_and_vec_path():
    # each partition is a whole number of 32-deep slabs
    assert list(tuning.k_partitions(np.array([16, 100, 512, 1000]), 4)) == [1, 4, 4, 4]
    assert list(tuning.k_partitions(np.array([64, 65, 96]), 16)) == [2, 3, 3]
    s = np.zeros(4, dtype=tuning.SHAPE_DTYPE)
    s[0] = (0, 0, 0, 100, 10, 64, 0)     # lda 100, ldb 64: both multiples of 4 floats
    s[1] = (0, 0, 0, 100, 10, 64, 2)     # lda 102
    s[2] = (1, 1, 1, 7, 9, 64, 0)        # FP64 lda = K = 64, ldb = N = 9 (odd)
    s[3] = (2, 1, 0, 8, 3, 12, 5)        # CGEMM3M: planes M = 8, K = 12
    assert list(tuning.vec_path(s)) == [True, False, False, True]


def test_library_features_equal_python_features():
    # the library's mlp_features (host code) against tuning.features on both
    # samplers' shapes, at the MI355X CU count and at another
    s = np.concatenate([tuning.sample_shapes(5, 3000), tuning.sample_shapes_aligned(6, 3000)])
    for cus in (256, 304):
        lib = tuning.library_features(s, cus=cus)
        py = tuning.features(s, cus=cus)
        assert np.array_equal(lib, py), np.argwhere(lib != py)[:5]


def test_sample_shapes_aligned():
    s = tuning.sample_shapes_aligned(3, 6000)
    assert np.array_equal(s, tuning.sample_shapes_aligned(3, 6000))
    for key in ("m", "n", "k"):
        v = s[key]
        assert v.min() >= 64 and v.max() <= 16384 and np.all(v % 64 == 0)
        # half the draws are powers of two; a quarter of the multiples of 64
        # (those <= 128 mostly) are too
        assert abs(np.mean((v & (v - 1)) == 0) - 0.63) < 0.03
    assert np.all(np.bincount(s["op"]) == 2000)
    assert abs(np.mean(s["pad"] == 0) - 0.75) < 0.03


def test_sample_shapes_distribution():
    s = tuning.sample_shapes(7, 6000)
    assert np.array_equal(s, tuning.sample_shapes(7, 6000))
    for key in ("m", "n", "k"):
        assert s[key].min() >= 16 and s[key].max() <= 16384
    assert np.all(np.bincount(s["op"]) == 2000)
    # 20% forced skinny K plus the 30% of log-uniform K that lands in [16, 128]
    assert abs(np.mean(s["k"] <= 128) - (0.2 + 0.8 * 0.3)) < 0.03
    # 20% forced small M or N plus log-uniform draws: P(min(M, N) <= 64) = 1 - 0.8^2
    assert abs(np.mean(np.minimum(s["m"], s["n"]) <= 64) - (0.2 + 0.8 * 0.36)) < 0.03
    assert 0.45 < np.mean(s["pad"] == 0) < 0.55 and s["pad"].max() <= 63
    big = s["k"] > 128
    assert abs(np.median(np.log2(s["k"][big])) - 10.5) < 0.2   # log-uniform over [2^7, 2^14]


def test_targets_and_regret():
    ms = np.full((2, 21), -1.0, dtype=np.float32)
    ms[0, [0, 1, 5, 20]] = [2.0, 1.0, 16.0, 2.0]
    ms[1, [3, 7, 20]] = [3.0, 6.0, 6.0]
    t, mask = tuning.targets(ms)
    assert t[0, 0] == 1.0 and t[0, 1] == 0.0 and t[0, 5] == 3.0     # log2(16) clipped at 3
    assert mask[0].sum() == 3 and mask[1].sum() == 2
    assert np.allclose(tuning.regret(ms, np.array([20, 3])), [1.0, 0.0])


def test_valid_mask_matches_rule():
    v = tuning.valid_mask(np.array([1, 0]), np.array([16384, 100]), np.array([16384, 100]), np.array([4096, 100]))
    assert v[0, 0] and not v[0, 1] and not v[0, 4]     # FP64 16384^2 partials: 2 GiB at split 2
    assert v[1, 0] and not v[1, 1] and not v[1, 2]     # K = 100 < 64 * 2
    assert v[1, 5] and not v[1, 6]                     # tile 1, split 1 / split 2


# ------------------------------------------------------------------ committed data and weights
DATA = tuning.DATA
HELD = os.path.join(DATA, "sweep_heldout.npz")
WEIGHTS = os.path.join(DATA, "picker.npz")


def _inc_params():
    txt = open(tuning.WEIGHTS_INC).read()
    m = re.search(r"kPickerWeights\[(\d+)\] = \{(.*?)\};", txt, re.S)
    vals = [v for v in re.split(r"[\s,]+", m.group(2)) if v]
    return np.array([float(v.rstrip("f")) for v in vals], dtype=np.float32)


needs_data = pytest.mark.skipif(not (os.path.exists(HELD) and os.path.exists(WEIGHTS)),
                                reason="sweep data / trained weights not generated yet")


@needs_data
def test_compiled_in_weights_equal_trained_weights():
    assert np.array_equal(_inc_params(), np.load(WEIGHTS)["params"])


ALIGNED = os.path.join(DATA, "sweep_aligned_heldout.npz")


@pytest.mark.skipif(not os.path.exists(ALIGNED), reason="aligned held-out sweep not generated yet")
def test_picker_regret_on_aligned_shapes():
    # power-of-two and multiple-of-64 sizes, where whole waves of tiles decide
    # the best configuration: 1 000 held-out shapes (seed 4)
    d = np.load(ALIGNED, allow_pickle=False)
    shapes, ms = d["shapes"], d["ms"]
    y = O.mlp_forward(np.load(WEIGHTS)["params"].astype(np.float64), tuning.features(shapes).astype(np.float64))
    choice = tuning.choose(y, tuning.valid_mask(shapes["op"], shapes["m"], shapes["n"], shapes["k"]))
    r_mlp = tuning.regret(ms, choice)
    r_heur = tuning.regret(ms, np.full(len(ms), 20))
    assert len(ms) >= 1000
    assert np.median(r_mlp) <= 0.01 and np.mean(r_mlp) <= 0.02 and np.percentile(r_mlp, 90) <= 0.05, (
        np.median(r_mlp), np.mean(r_mlp), np.percentile(r_mlp, 90))
    assert np.mean(r_mlp) < 0.25 * np.mean(r_heur)


@needs_data
def test_picker_regret_within_5pct_and_beats_heuristic():
    d = np.load(HELD, allow_pickle=False)
    shapes, ms = d["shapes"], d["ms"]
    y = O.mlp_forward(np.load(WEIGHTS)["params"].astype(np.float64), tuning.features(shapes).astype(np.float64))
    valid = tuning.valid_mask(shapes["op"], shapes["m"], shapes["n"], shapes["k"])
    choice = tuning.choose(y, valid)
    r_mlp = tuning.regret(ms, choice)
    r_heur = tuning.regret(ms, np.full(len(ms), 20))
    assert len(ms) >= 2000
    assert np.median(r_mlp) <= 0.05, np.median(r_mlp)
    assert np.mean(r_mlp) < np.mean(r_heur), (np.mean(r_mlp), np.mean(r_heur))
    for op in (0, 1, 2):
        sel = shapes["op"] == op
        assert np.median(r_mlp[sel]) <= 0.05, (op, np.median(r_mlp[sel]))
