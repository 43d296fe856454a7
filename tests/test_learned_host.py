"""Host-side logic of the learned warm start (no GPU): the training set's scale factors."""

import numpy as np

from ctrecon_amd.learned import running_std


def test_running_std_equals_the_std_of_the_concatenation():
    rng = np.random.default_rng(7)
    arrays = [rng.normal(0.3, 0.05, s).astype(np.float32)
              for s in [(20, 30, 40), (5, 7, 9), (1,), (64, 64)]]
    ref = float(np.std(np.concatenate([a.astype(np.float64).ravel() for a in arrays])))
    got = running_std(iter(arrays))  # a generator, as train() passes it
    assert abs(got - ref) <= 1e-12 * ref


def test_running_std_of_a_constant_is_zero():
    assert running_std([np.full((3, 4), 2.5, np.float32)] * 3) == 0.0
