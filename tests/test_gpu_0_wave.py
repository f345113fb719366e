"""Device self-test of the wave64 primitives (runs first among the GPU tests).

The counting kernel derives load addresses from DPP prefix sums, lane permutes
and ballot ranks; lstc_selftest compares them with host arithmetic on fixed
patterns.
"""
import pytest

pytestmark = pytest.mark.gpu


def test_wave_primitives(lstc):
    assert lstc.selftest() == 0
