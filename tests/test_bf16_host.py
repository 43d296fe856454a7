"""Known answers for the numpy BF16 emulation the BF16 GPU tests are built on."""

import numpy as np

from bf16 import quantize, round_bits, ulp, widen


def test_round_to_nearest_even_known_answers():
    vals = np.array([1.0, 1 + 2.0**-8, 1 + 3 * 2.0**-8, 1 + 2.0**-8 + 2.0**-20, -2.5,
                     3.14159265, 0.0, -0.0, np.finfo(np.float32).max], dtype=np.float32)
    # 1+2^-8 is a tie between 0x3F80 and 0x3F81: even 0x3F80; 1+3*2^-8 between
    # 0x3F81 and 0x3F82: even 0x3F82; just above a tie rounds up; pi = 0x40490FDB
    # keeps its upper half; FLT_MAX rounds to +inf
    want = [0x3F80, 0x3F80, 0x3F82, 0x3F81, 0xC020, 0x4049, 0x0000, 0x8000, 0x7F80]
    assert round_bits(vals).tolist() == want


def test_widen_is_exact_and_idempotent():
    rng = np.random.default_rng(0)
    a = rng.standard_normal(10000).astype(np.float32) * 1e3
    q = quantize(a)
    assert np.array_equal(quantize(q), q)
    assert np.array_equal(widen(round_bits(q)), q)
    # |a - q| <= half a BF16 ulp
    assert np.all(np.abs(a.astype(np.float64) - q) <= 0.5 * ulp(a) * (1 + 1e-12))
