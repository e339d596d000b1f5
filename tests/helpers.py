"""Shared helpers for the parity tests."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name), allow_pickle=False))


def rel_rms(approx, ref):
    """sqrt(sum |a - r|^2) / sqrt(sum |r|^2) over all reals."""
    a = np.asarray(approx, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    den = np.sqrt(np.sum(r * r))
    return float(np.sqrt(np.sum((a - r) ** 2)) / (den if den > 0 else 1.0))


def max_rel(approx, ref):
    """max |a - r| / max |r| (signed comparison of every real)."""
    a = np.asarray(approx, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    den = np.max(np.abs(r))
    return float(np.max(np.abs(a - r)) / (den if den > 0 else 1.0))


def xcorr_key(prec, matrix, gauss, casa, bs, ts):
    return "vis_%s_%s_%s_c%d_bs%d_ts%d" % (prec, "m" if matrix else "s",
                                           "g" if gauss else "p", casa, bs,
                                           ts)


XCORR_CASES = [(prec, matrix, gauss, casa, bs, ts)
               for prec in ("f", "d") for matrix in (0, 1)
               for gauss in (0, 1) for casa in (0, 1) for bs in (0, 1)
               for ts in (0, 1)]
