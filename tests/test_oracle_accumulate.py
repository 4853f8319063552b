"""The checker's accumulate_sample / mark_converged (src/main.rs:597-677)
pinned by the reference's own firefly tests (src/main.rs:1243-1280) and by
closed forms: Welford M2 equals the sample variance, the clamp lands a spike
exactly on max(clamp * mean luminance, floor), non-finite samples change
nothing, dark channels are exempt from the convergence test."""
import numpy as np

from oracle import oracle as O


def state(n):
    return dict(accum=np.zeros((n, 3), np.float32), samples=np.zeros(n, np.uint32),
                var_m2=np.zeros((n, 3), np.float32), converged=np.zeros(n, np.uint8))


def test_default_config_clamps_spike():
    s = state(1)
    for _ in range(4):
        O.accumulate_sample(np.zeros((1, 3), np.float32), s)
    O.accumulate_sample(np.full((1, 3), 100.0, np.float32), s)
    assert abs(s["accum"][0, 0] / s["samples"][0] - 5.0 / 5.0) < 1e-4


def test_zero_clamp_disables_firefly_clamp():
    s = state(1)
    for _ in range(4):
        O.accumulate_sample(np.zeros((1, 3), np.float32), s, firefly_clamp=0.0)
    O.accumulate_sample(np.full((1, 3), 100.0, np.float32), s, firefly_clamp=0.0)
    assert abs(s["accum"][0, 0] / s["samples"][0] - 20.0) < 1e-3


def test_welford_m2_is_the_sample_variance_and_the_sum_is_exact():
    g = np.random.default_rng(1)
    x = g.gamma(2.0, 0.5, size=(40, 200, 3)).astype(np.float32)
    s = state(200)
    for k in range(40):
        O.accumulate_sample(x[k], s, firefly_clamp=0.0)
    assert (s["samples"] == 40).all()
    assert np.allclose(s["accum"], x.astype(np.float64).sum(0), rtol=1e-5)
    assert np.allclose(s["var_m2"] / 39.0, x.astype(np.float64).var(0, ddof=1), rtol=1e-3, atol=1e-6)


def test_clamp_scales_a_spike_to_the_threshold():
    s = state(2)
    c = np.array([[0.5, 1.0, 2.0], [0.1, 0.1, 0.1]], np.float32)
    for _ in range(4):
        O.accumulate_sample(c, s)
    spike = np.array([[50.0, 20.0, 10.0], [0.2, 0.2, 0.2]], np.float32)
    before = s["accum"].copy()
    O.accumulate_sample(spike, s)
    lum = lambda v: 0.2126 * v[0] + 0.7152 * v[1] + 0.0722 * v[2]
    th = max(8.0 * lum(c[0]), 5.0)
    added = s["accum"][0] - before[0]
    assert abs(lum(added) - th) < 1e-4 * th
    assert np.allclose(added / added[0], spike[0] / spike[0][0], rtol=1e-5)   # chroma kept
    assert np.allclose(s["accum"][1] - before[1], spike[1])                   # below the floor: untouched


def test_non_finite_samples_and_converged_pixels_change_nothing():
    s = state(4)
    O.accumulate_sample(np.ones((4, 3), np.float32), s)
    s["converged"][3] = 1
    ref = {k: v.copy() for k, v in s.items()}
    bad = np.ones((4, 3), np.float32)
    bad[0, 0], bad[1, 1], bad[2, 2] = np.nan, np.inf, -np.inf
    bad[3] = 7.0
    O.accumulate_sample(bad, s)
    for k in ref:
        assert np.array_equal(ref[k], s[k]), k


def test_mark_converged_rules():
    s = state(6)
    n = 10
    s["samples"][:] = n
    # 0: low relative error; 1: high; 2: dark in every channel; 3: no samples;
    # 4: one bright channel with high error among dark ones; 5: already flagged
    s["accum"][0] = [5.0, 5.0, 5.0]
    s["var_m2"][0] = [0.9 * 9 * (0.05 * 0.5) ** 2 * n] * 3          # se / m = 0.05 * sqrt(0.9)
    s["accum"][1] = [5.0, 5.0, 5.0]
    s["var_m2"][1] = [1.1 * 9 * (0.05 * 0.5) ** 2 * n] * 3
    s["accum"][2] = [5e-4, 5e-4, 5e-4]
    s["var_m2"][2] = [100.0, 100.0, 100.0]
    s["samples"][3] = 0
    s["accum"][4] = [5e-4, 5.0, 5e-4]
    s["var_m2"][4] = [0.0, 100.0, 0.0]
    s["converged"][5] = 1
    cnt = O.mark_converged(s, 0.05)
    assert list(s["converged"]) == [1, 0, 1, 0, 0, 1] and cnt == 3
