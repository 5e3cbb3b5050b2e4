"""Pins the oracle's image geometry, voxel-ratio branch, store rounding and
word counters by statements derived without the oracle's own ray()
(tests/oracle_pins.py).  test_oracle_mutants.py shows that each pin fails
under the mistake it is meant to catch."""
import pytest

from tests import oracle_pins as P


@pytest.mark.parametrize("name", sorted(P.CHECKS))
def test_oracle_pin(name):
    P.CHECKS[name]()
