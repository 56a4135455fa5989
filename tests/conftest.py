import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: a CPU test of more than a few seconds")


class Lcg:
    """The reference tests' LCG (ybus_cuda/tests/common/mod.rs:20-32,
    ybus_sparse/tests/factorization_test.rs:15-34)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next(self) -> int:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & self.MASK
        return self.state >> 33

    def value(self) -> float:
        return (self.next() % 2001) / 1000.0 - 1.0

    def permutation(self, n: int):
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = self.next() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        return perm


@pytest.fixture
def lcg():
    return Lcg
