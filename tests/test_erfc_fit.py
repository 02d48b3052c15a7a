"""Error bound of the pair kernel's erfc (md_engine_amd/csrc/erfc_fit.h):
erfc(x) = exp(-x^2) t P(t), t = 1 / (1 + p x), with the constants as the
library ships them (FP32), against scipy's erfc: relative error on [0, 5]
below 5e-8 in FP64 evaluation and below 5e-7 in FP32 evaluation (about the
error of an FP32 erfc itself), absolute error below 5e-7 out to x = 10."""
import os
import re

import numpy as np
from scipy.special import erfc

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "md_engine_amd", "csrc",
                   "erfc_fit.h")


def constants():
    txt = open(HDR).read()
    p = float(re.search(r"kErfcP = ([-+0-9.e]+)f", txt).group(1))
    c = [float(v.rstrip("f")) for v in re.search(r"kErfcC\[8\] = \{([^}]*)\}", txt).group(1).split(",")]
    return np.float32(p), np.array(c, np.float32)


def evaluate(x, p, c, dtype):
    x = x.astype(dtype)
    one = dtype(1)
    t = (one / (one + dtype(p) * x)).astype(dtype)
    s = np.full_like(x, dtype(c[-1]))
    for k in range(len(c) - 2, -1, -1):
        s = (s * t + dtype(c[k])).astype(dtype)
    g = np.exp(-(x.astype(np.float64) ** 2)).astype(dtype)
    return (g * t * s).astype(np.float64)


def test_fp64_relative_error():
    p, c = constants()
    x = np.linspace(0.0, 5.0, 50001)
    rel = np.abs(evaluate(x, p, c, np.float64) / erfc(x) - 1)
    assert rel.max() < 5e-8, rel.max()


def test_fp32_relative_and_absolute_error():
    # against erfc at the same FP32 argument (the argument's own rounding
    # moves erfc by about 2 x^2 eps, beyond the approximation's doing)
    p, c = constants()
    x = np.linspace(0.0, 5.0, 50001).astype(np.float32).astype(np.float64)
    rel = np.abs(evaluate(x, p, c, np.float32) / erfc(x) - 1)
    assert rel.max() < 5e-7, rel.max()
    x = np.linspace(0.0, 10.0, 50001).astype(np.float32).astype(np.float64)
    assert np.abs(evaluate(x, p, c, np.float32) - erfc(x)).max() < 5e-7

