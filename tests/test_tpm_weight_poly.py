"""The TreePM short-range weight of the walk, w(x) = erfc(x/2) + x/sqrt(pi)
exp(-x^2/4) (x = r / r_s), is evaluated as exp(-x^2/4) Q(x) with Q a
degree-13 polynomial held at x = 8 beyond (blitzar_amd/csrc/k_walk.hip,
kTpmQ). This reads the committed coefficients from the source and checks,
in FP32 arithmetic (fused multiply-adds), that w errs by less than 2e-7
against the double-precision erfc everywhere, and that w(0) = 1."""
import math
import re
from pathlib import Path

import numpy as np

SRC = Path(__file__).resolve().parents[1] / "blitzar_amd" / "csrc" / "k_walk.hip"


def coefficients():
    text = SRC.read_text()
    m = re.search(r"kTpmQ\[14\]\s*=\s*\{([^}]*)\}", text)
    assert m, "kTpmQ not found"
    vals = [float(v.strip().rstrip("f")) for v in m.group(1).split(",") if v.strip()]
    assert len(vals) == 14
    return np.array(vals, np.float32)


def f32(v):
    return np.asarray(v, np.float64).astype(np.float32).astype(np.float64)


def weight_fp32(x):
    c = coefficients().astype(np.float64)
    x = f32(x)
    xc = np.minimum(x, 8.0)
    q = np.full_like(x, c[13])
    for k in range(12, -1, -1):
        q = f32(q * xc + c[k])  # fused: exact product and sum, one rounding
    e = f32(np.exp2(f32(-f32(x * x) * f32(0.3606737554))))
    return f32(q * e)


def test_weight_polynomial_matches_erfc():
    x = np.linspace(0.0, 12.0, 240001)
    ref = np.array([math.erfc(v / 2) + v / math.sqrt(math.pi) * math.exp(-v * v / 4) for v in x])
    w = weight_fp32(x)
    err = np.abs(w - ref)
    assert err.max() < 2e-7, (err.max(), x[err.argmax()])
    assert weight_fp32(np.array([0.0]))[0] == 1.0


def test_weight_is_monotone_and_small_beyond_the_cutoff():
    x = np.linspace(0.0, 12.0, 24001)
    w = weight_fp32(x)
    assert np.all(np.diff(w) <= 1e-7)      # w decreases (to FP32 rounding)
    assert w[np.searchsorted(x, 4.5)] < 0.02  # the usual cutoff, 4.5 r_s
