"""Regression test of the hi + lo split of the factorised correlator's
16-bit operands (oskar_cross_correlate_factorised, FP16x3 and BF16x3).

Each W operand K_q conj(C) is an FP32 product split into hi + lo 16-bit
parts.  Both parts must come from the one rounded FP32 value: a split that
rounds the exact product to 16 bits for one copy of hi and the rounded
product for another misses by one 16-bit unit wherever the rounded product
lies on a 16-bit rounding tie (about one value in 2^13), a 2^-11 error on
that product.  The sky here makes such a fault visible: 64 sources 10^4
times brighter than the other 4,032 dominate every visibility, so their
W values on 512 stations (about 260,000 values, some 30 of them on ties)
would carry faults of about 2e-5 of the largest visibility.  Checked
against FP64 visibilities of the same FP32-rounded inputs on every
baseline: the largest error of any baseline, relative to the largest
visibility, below 5e-6 for FP16x3 (measured 1.5e-6; 2.5e-5 with the faulty
split) and 1e-5 for BF16x3, whose 8-bit parts leave 4.3e-6.
"""
import numpy as np
import pytest

from fact_cases import make_case
from fp64_reference import direct_fp64_vis, f32
from gpu_util import gpu_factorised

pytestmark = pytest.mark.gpu
FREQ = 150e6


def bright_case():
    c = make_case(91, 512, 4096, spread=20000.0, pol=False)
    rng = np.random.default_rng(92)
    c["I"] = np.ones(4096)
    c["I"][rng.choice(4096, 64, replace=False)] = 1e4
    return c


def worst_baseline_error(got, ref):
    d = np.abs(np.asarray(got, np.float64) - ref).reshape(-1, 8).max(1)
    return float(d.max() / np.abs(ref).max())


@pytest.fixture(scope="module")
def case_and_ref(hip):
    c = bright_case()
    ref = direct_fp64_vis(*(f32(c[k]) for k in ("u", "v", "w", "l", "m",
                                                 "n", "I")), FREQ,
                          beam=f32(c["E"]))
    return c, ref


@pytest.mark.parametrize("mode", [0, 1], ids=["fp16x3", "bf16x3"])
def test_split_faults_absent(case_and_ref, mode):
    c, ref = case_and_ref
    got = gpu_factorised("f", mode, 0, c["E"].ravel(), c["I"], c["Q"],
                         c["U"], c["V"], c["l"], c["m"], c["n"], c["u"],
                         c["v"], c["w"], FREQ)
    err = worst_baseline_error(got, ref)
    print("split test %s: worst baseline |dV| / max |V64| %.3e"
          % (("fp16x3", "bf16x3")[mode], err), flush=True)
    assert err < (5e-6, 1e-5)[mode]
