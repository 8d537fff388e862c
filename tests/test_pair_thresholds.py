"""The packed domino pair kernel's integer form of the two-periodic
heat-bath test (dominokernelCFTPtwoperiodic.cl:15-44, restated by
domino_move): with rd = u 2^-23, u = b >> 9 the draw's grid value,

    rd < 0.5:  flip 3 -> 12 iff ratio1 > 2 rd        (float compare)
    else:      flip 12 -> 3 iff ratio2 > 2 (rd - 0.5)

is "(u mod 2^22) < ceil(min(ratio, 1) 2^22)" for the branch's ratio.  Checked
for every one of the 2^23 grid values and several weight pairs, the
ratios formed as the kernel forms them (float division of the squared
weights rounded to float, DominoWeights)."""
import math

import numpy as np
import pytest


def _k(ratio):
    """ceil(min(ratio, 1) * 2^22), as the kernel computes it in float."""
    r = np.float32(ratio)
    if r >= np.float32(1.0):
        return 1 << 22
    return int(math.ceil(float(r * np.float32(4194304.0))))


@pytest.mark.parametrize("a,b", [(1.0, 0.5), (0.7, 1.0), (1.0, 0.7), (0.9, 1.3), (1.0, 1.0),
                                 (0.31, 2.7)])
def test_two_periodic_integer_thresholds_equal_the_float_rule(a, b):
    wa2 = np.float32(a * a)  # (float)(a*a), the product in double as the reference's literals
    wb2 = np.float32(b * b)
    u = np.arange(1 << 23, dtype=np.uint32)
    rd = (u.astype(np.float32) * np.float32(2.0 ** -23)).astype(np.float32)
    lo = rd < np.float32(0.5)
    th = np.where(lo, rd * np.float32(2.0), (rd - np.float32(0.5)) * np.float32(2.0))
    assert th.dtype == np.float32
    u2 = u & np.uint32(0x3FFFFF)
    for r1, r2 in ((wa2 / wb2, wb2 / wa2), (wb2 / wa2, wa2 / wb2)):
        r1, r2 = np.float32(r1), np.float32(r2)
        want = np.where(lo, r1 > th, r2 > th)
        k = np.where(lo, _k(r1), _k(r2)).astype(np.int64)
        got = u2.astype(np.int64) < k
        assert np.array_equal(got, want), (a, b, int(np.argmax(got != want)))
