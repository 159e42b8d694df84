"""CPU check that the stated measurement order is observable in the GPU test's
cases: with the butterfly's offsets taken in ascending order (1, 2, ..., 32)
instead, the collapsed state of some of those cases rounds differently, so a
device that paired lanes in another order would fail the bitwise test."""
import numpy as np

from tests.test_gpu_reduction_order import measured_seq, random_case


def test_another_pairing_order_changes_the_collapsed_bits():
    for n in (6, 8, 10, 12, 14):
        rng = np.random.default_rng(500 + n)   # the GPU test's cases
        differ = 0
        for _ in range(12):
            re, im, qs, rds = random_case(rng, n)
            _, r_a, i_a = measured_seq(re, im, qs, rds)
            _, r_b, i_b = measured_seq(re, im, qs, rds, order=(1, 2, 4, 8, 16, 32))
            differ += int(np.any(r_a.view(np.uint64) != r_b.view(np.uint64)) or
                          np.any(i_a.view(np.uint64) != i_b.view(np.uint64)))
        assert differ >= 2, (n, differ)
