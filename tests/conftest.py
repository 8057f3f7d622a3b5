import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


Q = 1125899906840833  # largest prime < 2^50 with q = 1 mod 128
D = 64
NR = 5  # smallest quadratic non-residue mod Q
# The reference's own rings (protocol.rs:28-47, the A0/A1/A2 parameter sets):
# (degree N, MOD_Q, smallest quadratic non-residue mod MOD_Q)
REF_RINGS = [(128, 1125899904679937, 3), (256, 1125899906840833, 5),
             (512, 1125899906822657, 3)]


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle
