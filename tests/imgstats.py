"""Per-pixel statistics for the image-level parity gate.

A "frame" is the mean of S consecutive samples per pixel; K frames per side
give each pixel channel a mean and a sample variance (float64).  Two unbiased
renderers of the same image then differ per channel by a Welch z-score

    z = |m_a - m_b| / sqrt(v_a / K_a + v_b / K_b),

which for frame means that are close to normal (S samples each) follows a
t-distribution with about K_a + K_b - 2 degrees of freedom: about 0.3 % of
channels beyond 3 sigma and, for practical purposes, none beyond 6 sigma.
The gate is the one the blueprint states: at most 0.5 % of channels beyond
3 sigma and none beyond 6 sigma.  The standard error carries a floor of
1e-5 of the mean (float rounding, see REL_FLOOR).  Channels where both sides
never vary (black background, a wall no path reaches) count as agreeing when
their means are equal to 1e-6 relative, and as failing otherwise.
"""
import numpy as np

GATE_3 = 0.005


class FrameStats:
    """Running per-pixel mean / M2 over frames (Welford, float64)."""

    def __init__(self):
        self.k = 0
        self.mean = None
        self.m2 = None

    def add(self, frame):
        f = np.asarray(frame, np.float64)
        if self.mean is None:
            self.mean = np.zeros_like(f)
            self.m2 = np.zeros_like(f)
        self.k += 1
        d = f - self.mean
        self.mean += d / self.k
        self.m2 += d * (f - self.mean)
        return self

    @property
    def var(self):
        return self.m2 / max(self.k - 1, 1)


def frames(render, K, S):
    """FrameStats of K frames; render(first_sample, S) returns the SUM of S
    passes starting at sample index first_sample (an (H, W, 3) array)."""
    st = FrameStats()
    for k in range(K):
        st.add(np.asarray(render(k * S, S), np.float64) / S)
    return st


# Differences below this fraction of the mean are float rounding, not
# sampling: a pixel that sees one constant texel has next to no variance, and
# the device and the oracle round its bilinear lookup 1 ulp apart.  A 4,096-
# sample f32 mean carries about this much accumulation error itself.
REL_FLOOR = 1e-5


def z_scores(a, b):
    se = np.sqrt(a.var / a.k + b.var / b.k + (REL_FLOOR * np.maximum(np.abs(a.mean), np.abs(b.mean))) ** 2)
    diff = np.abs(a.mean - b.mean)
    z = diff / np.maximum(se, 1e-300)
    const = (a.var == 0) & (b.var == 0)
    same = diff <= 1e-6 * np.maximum(np.abs(a.mean), 1.0)
    return np.where(const, np.where(same, 0.0, np.inf), z)


def gate(a, b, frac3=GATE_3):
    """Assert the 3-sigma gate; returns the z-scores."""
    z = z_scores(a, b)
    f3, n6 = float((z > 3).mean()), int((z > 6).sum())
    assert f3 <= frac3 and n6 == 0, f"{f3:.4f} of channels beyond 3 sigma (gate {frac3}), {n6} beyond 6 sigma"
    return z
