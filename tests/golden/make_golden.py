"""Writes tests/golden/tiny_dd3.npz: oracle (FP64) outputs on the tiny and ragged
geometries, used as regression fixtures by the CPU and GPU tests.

    python tests/golden/make_golden.py

Inputs are seeded; outputs come from the FP64 oracle (integral-image transcription
of the reference arithmetic, oracle/dd3_oracle.cpp).  The reference tree holds no
output vectors for this path, so these fixtures pin the oracle against later
changes rather than against a run of the reference.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from cases import ragged, sinogram, small, tiny, volume  # noqa: E402
from oracle import oracle as O  # noqa: E402
from ctrecon_amd import phantom as P  # noqa: E402


def main():
    out = {}
    for name, fn in (("tiny", tiny), ("ragged", ragged)):
        cfg, g = fn()
        vol = volume(g, seed=11)
        y = sinogram(g, seed=12)
        out[f"{name}_vol"] = vol
        out[f"{name}_sino"] = y
        out[f"{name}_fp"] = O.fp(g, vol, "sat")
        out[f"{name}_bp_ref"] = O.bp(g, y, "sat")
        out[f"{name}_bp_adj"] = O.bp(g, y, "adjoint")
    cfg, g = small()
    sl = P.shepp_logan(g, seed=0)
    out["small_shepp_logan_sum"] = np.array(sl.astype(np.float64).sum())
    out["small_shepp_logan_nnz"] = np.array((sl > 0).sum())
    np.savez_compressed(os.path.join(HERE, "tiny_dd3.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
