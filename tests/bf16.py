"""BF16 rounding emulated in numpy (round to nearest even, finite values), used by the
tests to build the exact inputs and expected outputs of the BF16-storage kernels."""

import numpy as np


def round_bits(a):
    """float32 -> BF16 bit patterns (uint16), round to nearest, ties to even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r & 0xFFFF).astype(np.uint16)


def widen(h):
    """BF16 bit patterns -> float32 (exact)."""
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def quantize(a):
    """The float32 value of round_bits(a)."""
    return widen(round_bits(a))


def ulp(a):
    """Spacing of BF16 values at |a| (8 significant bits): 2^(floor(log2|a|) - 7)."""
    a = np.abs(np.asarray(a, dtype=np.float64))
    e = np.floor(np.log2(np.maximum(a, np.finfo(np.float32).tiny)))
    return np.exp2(e - 7)
