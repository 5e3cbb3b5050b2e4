"""Mutation tests of the oracle's pins: each build of the oracle with one
deliberate mistake planted (oracle/xo_trace_body.h, XO_MUTANT=k) must fail
every pin named for it in oracle_pins.MUTANT_CATCHES.  The mistakes are
planted in the CPU oracle only; nothing here touches a GPU."""
import pytest

from oracle import xo
from tests import oracle_pins as P


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("mutant", sorted(P.MUTANT_CATCHES))
def test_pins_catch_planted_mistake(mutant, mutant_dir):
    path = xo.build_mutant(mutant, mutant_dir)
    with xo.use_library(path):
        for name in P.MUTANT_CATCHES[mutant]:
            with pytest.raises(AssertionError):
                P.CHECKS[name]()
    # the faithful oracle is back in place and passes the same pins
    for name in P.MUTANT_CATCHES[mutant]:
        P.CHECKS[name]()
