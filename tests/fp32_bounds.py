"""Parity bars for an FP32 Joseph forward whose sums agree with elsa's to FP32 rounding.

The plane-major forward interpolates each sample in lerp form instead of elsa's weight
products (JosephsMethod.cpp:60-97), so it is not bit-equal to the oracle's FP32 restatement.
It is checked against two oracle results on the same FP32 geometry (same sample points, floors
and fractions):
  ref32  the oracle's FP32 forward (elsa's association, (value * len) * weight per tap),
  exact  the exact bilinear interpolation at those points in FP64 (MODE_EXACT_WEIGHTS, acc64:
         weights 1 - c and their products formed in FP64), with mag the same for |value|.
Bars: rel. L2 to ref32 <= 1e-6; L2 error to exact no larger than ref32's own (times 1.05);
every element within four times the worst relative error (to mag) that ref32 shows, and at
least within 2^-22 of mag. A wrong tap, fraction or axis moves elements by O(value), far outside.
"""
import numpy as np

import oracle


def oracle_forward_refs(g, x):
    fn = oracle.joseph
    x = np.asarray(x, np.float32)
    ref32 = fn(g, x, mode=oracle.MODE_DIRECT)
    exact = fn(g, x.astype(np.float64), mode=oracle.MODE_EXACT_WEIGHTS, acc64=True)
    mag = fn(g, np.abs(x).astype(np.float64), mode=oracle.MODE_EXACT_WEIGHTS, acc64=True)
    return ref32, exact, mag


def assert_forward_close(out, ref32, exact, mag, what=""):
    out = np.asarray(out, np.float64)
    ref = np.asarray(ref32, np.float64)
    rel = np.linalg.norm(out - ref) / max(np.linalg.norm(ref), 1e-300)
    assert rel <= 1e-6, (what, rel)
    e_out = np.linalg.norm(out - exact)
    e_ref = np.linalg.norm(ref - exact)
    assert e_out <= 1.05 * e_ref + 1e-12 * np.linalg.norm(exact), (what, e_out, e_ref)
    with np.errstate(divide="ignore", invalid="ignore"):
        worst = np.max(np.where(mag > 0, np.abs(ref - exact) / mag, 0.0))
    bound = 4.0 * max(worst, 2.0 ** -24) * mag
    bad = np.abs(out - exact) > bound
    assert not np.any(bad), (what, int(bad.sum()), float(np.max(np.abs(out - exact) - bound)))


def check_forward(out, g, x, what=""):
    """assert_forward_close against the oracle's results for the Joseph forward of x."""
    assert_forward_close(out, *oracle_forward_refs(g, x), what=what)
