"""Regression fixtures (tests/golden/tiny_dd3.npz, written by make_golden.py)."""

import os

import numpy as np
import pytest

from conftest import rel_l2
from cases import ragged, small, tiny

from ctrecon_amd import phantom as P

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "tiny_dd3.npz"))
CASES = {"tiny": tiny, "ragged": ragged}


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(name, oracle):
    cfg, g = CASES[name]()
    vol, y = GOLD[f"{name}_vol"], GOLD[f"{name}_sino"]
    assert rel_l2(oracle.fp(g, vol, "sat"), GOLD[f"{name}_fp"]) < 1e-13
    assert rel_l2(oracle.bp(g, y, "sat"), GOLD[f"{name}_bp_ref"]) < 1e-13
    assert rel_l2(oracle.bp(g, y, "adjoint"), GOLD[f"{name}_bp_adj"]) < 1e-13


def test_phantom_checksum():
    cfg, g = small()
    sl = P.shepp_logan(g, seed=0)
    assert abs(sl.astype(np.float64).sum() - float(GOLD["small_shepp_logan_sum"])) < 1e-6
    assert int((sl > 0).sum()) == int(GOLD["small_shepp_logan_nnz"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_gpu_matches_golden(name, gpu_plan_factory):
    print(f"[golden gpu] case={name}", flush=True)
    cfg, g = CASES[name]()
    plan = gpu_plan_factory(g)
    vol, y = GOLD[f"{name}_vol"], GOLD[f"{name}_sino"]
    assert rel_l2(plan.forward(vol), GOLD[f"{name}_fp"]) <= 1e-5
    assert rel_l2(plan.back(y, kind="ref"), GOLD[f"{name}_bp_ref"]) <= 1e-5
    assert rel_l2(plan.back(y, kind="matched"), GOLD[f"{name}_bp_adj"]) <= 1e-5
